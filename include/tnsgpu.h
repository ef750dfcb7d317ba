/*
 * tnsgpu.h -- C ABI of libtnsgpu.so, the MI355X (gfx950) block-sparse tensor
 * engine.  Plain pointers and sizes only; device buffers belong to the caller
 * (the Python package allocates them with torch), the library never hands out
 * device memory.  Every function returns 0 on success; on failure it returns
 * a non-zero code and tns_last_error() gives a thread-local message.
 *
 * Reference interfaces replaced (paths relative to the TNSP repository root):
 *   shapes       TAT::Tensor(names, edges)          PyTAT/PyTAT.hpp:774-779,
 *                                                   TAT/include/TAT/structure/core.hpp:152-191
 *   edge op      Tensor::edge_operator               PyTAT/PyTAT.hpp:1000-1015,
 *                                                   TAT/include/TAT/implement/edge_operator.hpp:35-692
 *   contract     Tensor::contract                    PyTAT/PyTAT.hpp:988-1004,
 *                                                   TAT/include/TAT/implement/contract.hpp:307-620
 *   svd          Tensor::svd                         PyTAT/PyTAT.hpp:1031-1058,
 *                                                   TAT/include/TAT/implement/svd.hpp:260-538
 *   qr           Tensor::qr                          PyTAT/PyTAT.hpp:1059-1074,
 *                                                   TAT/include/TAT/implement/qr.hpp:310-508
 *   conjugate    Tensor::conjugate                   PyTAT/PyTAT.hpp:1029,
 *                                                   TAT/include/TAT/implement/conjugate.hpp:32-120
 *   amplitude    Configuration.hole(()) / replace    tetragono/tetragono/auxiliaries/single_layer_auxiliaries.py:463-498,
 *                                                   500-720
 *
 * Names are int32 ids (the caller interns strings).  Symmetry kinds:
 *   0 No, 1 BoseZ2, 2 BoseU1, 3 FermiU1, 4 FermiZ2, 5 FermiU1BoseZ2,
 *   6 FermiU1BoseU1, 7 FermiU1FermiU1.
 * A symmetry value is 2 int32 (unused components 0; Z2 as 0/1).
 * dtype: 0 = float64, 1 = float32.
 */
#ifndef TNSGPU_H
#define TNSGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tns_shape tns_shape;
typedef struct tns_plan tns_plan;

/* library */
const char* tns_last_error(void);
const char* tns_version(void);
int tns_device_count(int* count);
/* Jacobi sweep limit of the SVD kernels (default 60, range 1..1000).  A
 * sector that has not converged within the limit makes the SVD fail with an
 * error, as a LAPACK gesvd failure (info > 0) makes TAT raise
 * (TAT/include/TAT/implement/svd.hpp:125-139 with TAT_ERROR_BITS); lowering
 * the limit forces that path in tests.  Process-wide. */
int tns_set_svd_max_sweeps(int sweeps);

/* shapes (TAT::Tensor(names, edges): block layout of core.hpp:152-191) */
int tns_shape_create(int kind, int rank, const int32_t* names, const int32_t* nseg, const int32_t* syms,
                     const int64_t* dims, const uint8_t* arrows, tns_shape** out);
int tns_shape_free(tns_shape* s);
int tns_shape_counts(const tns_shape* s, int32_t* kind, int32_t* rank, int32_t* total_segments, int64_t* nblocks,
                     int64_t* size);
int tns_shape_export(const tns_shape* s, int32_t* names, int32_t* nseg, int32_t* syms, int64_t* dims, uint8_t* arrows);
/* block offsets in storage order: positions (nblocks x rank) and offsets (nblocks) */
int tns_shape_blocks(const tns_shape* s, int64_t* positions, int64_t* offsets);
int tns_shape_rename(const tns_shape* s, int n, const int32_t* old_names, const int32_t* new_names, tns_shape** out);

/* plans.  Output shapes are returned as new handles owned by the caller. */
int tns_plan_edge_operator(const tns_shape* src, int n_split, const int32_t* split_old, const int32_t* split_count,
                           const int32_t* split_names, const tns_shape* split_edges, int n_rev, const int32_t* rev,
                           int n_merge, const int32_t* merge_new, const int32_t* merge_count, const int32_t* merge_old,
                           int n_new, const int32_t* new_names, int apply_parity, int n_ex_split,
                           const int32_t* ex_split, int n_ex_rb, const int32_t* ex_rb, int n_ex_ra,
                           const int32_t* ex_ra, int n_ex_merge, const int32_t* ex_merge, tns_plan** out);
int tns_plan_contract(const tns_shape* a, const tns_shape* b, int npairs, const int32_t* pairs, tns_plan** out);
int tns_plan_svd(const tns_shape* t, int nfree, const int32_t* free_u, int32_t cu, int32_t cv, int32_t su,
                 int32_t sv, tns_plan** out);
int tns_plan_qr(const tns_shape* t, char direction, int nfree, const int32_t* free_names, int32_t cq, int32_t cr,
                tns_plan** out);
int tns_plan_conjugate(const tns_shape* t, int trivial_metric, tns_plan** out);
int tns_plan_free(tns_plan* p);

/* number of outputs and each output shape (for svd: after tns_svd_factor) */
int tns_plan_num_outputs(const tns_plan* p, int* n);
int tns_plan_output_shape(const tns_plan* p, int which, tns_shape** out);
/* device workspace the caller must provide to run the plan */
int tns_plan_workspace(const tns_plan* p, int dtype, int64_t* bytes);
/* flop count of the plan (sector GEMMs, QR/SVD counted as in LAPACK) */
int tns_plan_flops(const tns_plan* p, int64_t* flops);

/* execute on a HIP stream (NULL = default stream) */
int tns_plan_run(tns_plan* p, int dtype, const void* const* inputs, void* const* outputs, void* workspace,
                 void* stream);
/* tns_plan_run without any device allocation.  No counterpart in the
 * reference (TAT allocates per op); it is for callers that cannot allocate:
 * code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  The plan's workspace (at most 64 MiB) and its task records (at
 * most 4 MiB) come from one static device array the library holds on each GPU
 * until the process ends, instead of a caller workspace and the library's
 * growable staging.  A plan needing more workspace fails with an error
 * (before any device call); svd plans are refused (tns_svd_factor writes into
 * the caller's workspace).  Nothing in the array needs to start at zero.
 * Every call on one GPU shares that GPU's array: calls on one stream are
 * ordered, calls that overlap on different streams of one GPU race. */
int tns_plan_run_noalloc(tns_plan* p, int dtype, const void* const* inputs, void* const* outputs, void* stream);
/* svd phase 1: merge + per-sector SVD, then (synchronising) the TAT truncation
 * rule (svd.hpp:430-481): remain_cut < 0 means no count limit, relative_cut
 * <= 0 means none.  Output shapes are known afterwards; tns_plan_run then
 * writes U, S, V. */
int tns_svd_factor(tns_plan* p, int dtype, const void* input, void* workspace, int64_t remain_cut,
                   double relative_cut, void* stream);

/* the same op for n tensors: planning spread over `threads` host threads,
 * execution as one launch per kernel kind for all n plans (the reference has
 * no batched form; each plan is the one tns_plan_contract / _qr / _svd
 * builds; tns_plan_conjugate_many: Tensor::conjugate, conjugate.hpp:32-120).
 * tns_plans_run: plans of one type; inputs n x (2 for contract, 1 for qr /
 * svd / conjugate), outputs n x (1 contract / conjugate, 2 qr, 3 svd),
 * workspaces n (each tns_plan_workspace bytes).  Svd plans need
 * tns_svd_factor_many first, which applies one truncation rule to every
 * tensor. */
int tns_plan_contract_many(int n, const tns_shape* const* a, const tns_shape* const* b, int npairs,
                           const int32_t* pairs, int threads, tns_plan** out);
int tns_plan_qr_many(int n, const tns_shape* const* t, char direction, int nfree, const int32_t* free_names,
                     int32_t cq, int32_t cr, int threads, tns_plan** out);
int tns_plan_svd_many(int n, const tns_shape* const* t, int nfree, const int32_t* free_u, int32_t cu, int32_t cv,
                      int32_t su, int32_t sv, int threads, tns_plan** out);
int tns_plan_conjugate_many(int n, const tns_shape* const* t, int trivial_metric, int threads, tns_plan** out);
int tns_plans_run(int n, tns_plan* const* plans, int dtype, const void* const* inputs, void* const* outputs,
                  void* const* workspaces, void* stream);
int tns_svd_factor_many(int n, tns_plan* const* plans, int dtype, const void* const* inputs,
                        void* const* workspaces, int64_t remain_cut, double relative_cut, int threads, void* stream);
/* item-wise helpers of the batched ops (no reference counterpart: the
 * reference handles one tensor per call): free n handles; edge_rename of n
 * shapes with one mapping (tensor.hpp:760-784); output shape `which` of n
 * plans with its element count; workspace bytes of n plans. */
int tns_shapes_free(int n, tns_shape* const* shapes);
int tns_plans_free(int n, tns_plan* const* plans);
int tns_shape_rename_many(int n, const tns_shape* const* shapes, int nren, const int32_t* old_names,
                          const int32_t* new_names, tns_shape** out);
int tns_plans_output_shapes(int n, tns_plan* const* plans, int which, tns_shape** out, int64_t* sizes);
int tns_plans_workspace(int n, tns_plan* const* plans, int dtype, int64_t* bytes);

/* batched amplitude engine: psi(s) of many configurations of one PEPS in
 * lockstep (tetragono's single-layer boundary MPS,
 * tetragono/tetragono/auxiliaries/single_layer_auxiliaries.py:463-498,
 * 710-720).  sites: L1*L2 raw site shapes (with the physical edge) and their
 * device data; name_ids: P0 T U D L R L1 L2 L3 R1 R2 P Q then the L1*L2
 * projected physical names P_{l1}_{l2}_0 row-major.  occupations: nconf x
 * L1 x L2 indices along the physical edge.  psi: host output (nconf).
 * The engine owns its device arenas until tns_engine_free (tns_engine_trim
 * frees the working ones early). */
typedef struct tns_engine tns_engine;
int tns_engine_create(int kind, int L1, int L2, const tns_shape* const* sites, const void* const* site_data,
                      int dtype, const int32_t* name_ids, int n_names, int threads, tns_engine** out);
int tns_engine_set_chunk(tns_engine* e, int64_t chunk);
/* lanes: independent lockstep pipelines over the same PEPS, each on its own
 * host thread and HIP stream; configurations are split over them so one
 * lane's host planning and decomposition tails overlap the others' kernels
 * (1 <= lanes <= 32) */
int tns_engine_set_lanes(tns_engine* e, int lanes);
/* precision of the contraction GEMMs of a float64 engine: 0 fp64, 1 fp32,
 * 2 split-bf16 x3, 3 split-bf16 x6 (fp32 accumulation); operands are
 * converted on the way into the GEMM and the product back to float64, QR, SVD
 * and all stored tensors stay float64.  No reference counterpart (TAT
 * computes in the tensor's own scalar type). */
int tns_engine_set_precision(tns_engine* e, int precision);
/* GEMM tile autotuner (no reference counterpart): evaluates the given
 * configurations on lane 0 and, at every contraction of that evaluation,
 * times each GEMM tile variant (0: 32x32, 1: 64x64, 2: 16x16) on the
 * contraction's own sector GEMMs for each size class (max(m, n) <= 16,
 * <= 48, larger); the variant with the least total time per class is used
 * by every lane from then on.  table (3) gets the choice, ms (9,
 * class-major) the summed kernel times; either may be NULL. */
int tns_engine_autotune(tns_engine* e, int64_t nconf, const int32_t* occupations, int64_t remain_cut,
                        double relative_cut, int32_t* table, double* ms, void* stream);
int tns_engine_evaluate(tns_engine* e, int64_t nconf, const int32_t* occupations, int64_t remain_cut,
                        double relative_cut, double* psi, void* stream);
/* values: plan_s exec_s total_s gemm_flops qr_flops svd_flops perm_elems
 * launches device_bytes configs bind_s sync_s (n <= 12); plan_s is the
 * parallel per-configuration planning, bind_s the serial binding and task
 * packing, sync_s the waits for singular values (summed over lanes) */
int tns_engine_stats(const tns_engine* e, double* values, int n);
int tns_engine_reset_stats(tns_engine* e);
/* per-kernel device timing with HIP events on the launch stream: permute,
 * fill, gemm, qr (LDS / unblocked), svd (LDS / global), diag, qr_blocked,
 * svd_reg8, svd_reg16, svd_reg24; per kernel: launches, total ms,
 * algorithmic work (bytes for permute/fill/diag, flops for the others) */
int tns_engine_set_profile(tns_engine* e, int on);
/* the caller changed the site tensors' values in place (same buffers and
 * block structure), e.g. a gradient step (sampling_lattice/lattice.py:936-948):
 * the projected sites are rebuilt on the next evaluation */
int tns_engine_refresh_sites(tns_engine* e);
/* frees the engine's working device memory (boundary-MPS rows, temporaries,
 * singular values; the projected sites stay): it grows again on the next
 * evaluation.  No counterpart in the reference (TAT allocates per tensor);
 * for callers that share the GPU's memory with other work between
 * evaluations */
int tns_engine_trim(tns_engine* e);
int tns_engine_kernel_stats(const tns_engine* e, double* values, int n);
int tns_engine_free(tns_engine* e);

/* hop-proposal network for Metropolis sampling (no reference counterpart:
 * the reference proposes hops bond by bond, sampling.py:117-154).  NCHW
 * float32 activations, bf16 weights (uint16 bit patterns) [cout][cin][3][3],
 * float32 accumulation, zero padding.  relu_y: the layer's ReLU output whose
 * mask gates dy (NULL = no ReLU).  Hops are d * H * W + site with d = 0 up,
 * 1 down, 2 left, 3 right; a hop is valid when the site is filled and its
 * neighbour exists and is empty. */
int tns_conv3x3_forward(const float* x, const uint16_t* w, const float* bias, float* y, int batch, int cin, int cout,
                        int height, int width, int relu, void* stream);
int tns_conv3x3_backward_data(const float* dy, const float* relu_y, const uint16_t* w, float* dx, int batch, int cin,
                              int cout, int height, int width, void* stream);
int tns_conv3x3_backward_weight_workspace(int batch, int cin, int cout, int64_t* bytes);
int tns_conv3x3_backward_weight(const float* dy, const float* relu_y, const float* x, float* dw, float* dbias,
                                int batch, int cin, int cout, int height, int width, void* workspace, void* stream);
/* masked softmax over valid hops: draw one per configuration (inverse CDF
 * of uniforms in [0, 1)) or score the given ones; log_q = log q(hop) */
int tns_hop_sample(const float* logits, const int8_t* occupations, const float* uniforms, int32_t* hops,
                   float* log_q, int batch, int height, int width, void* stream);
int tns_hop_log_prob(const float* logits, const int8_t* occupations, int32_t* hops, float* log_q, int batch,
                     int height, int width, void* stream);
/* loss_b = (logit[b, hop_b] - target_b)^2, dlogits = d mean(loss) / d logits */
int tns_hop_regression(const float* logits, const int32_t* hops, const float* targets, float* dlogits, float* loss,
                       int batch, int n, void* stream);
int tns_adam_step(float* w, float* m, float* v, const float* grad, uint16_t* w_bf16, int n, float lr, float beta1,
                  float beta2, float eps, int step, void* stream);
int tns_to_bf16(const float* w, uint16_t* w_bf16, int n, void* stream);
int tns_occupations_to_input(const int8_t* occupations, float* x, int64_t n, void* stream);

/* planner introspection used by the CPU test-suite: element counts of the
 * plan's logical buffers (-1 = caller-owned input), the sector table of an
 * svd plan, and the truncation rule applied to host-computed singular
 * values (the GPU path applies the same rule inside tns_svd_factor) */
int tns_plan_buffers(const tns_plan* p, int64_t* sizes, int* n);
int tns_svd_sectors(const tns_plan* p, int64_t* k, int64_t* off, int32_t* syms, int* n);
int tns_svd_truncate(tns_plan* p, const double* svals, int64_t remain_cut, double relative_cut);

/* planner introspection used by the CPU test-suite: task records of a stage
 * (stage ids and record layouts are documented in capi.cpp) */
int tns_plan_stage_count(const tns_plan* p, int stage, int64_t* n);
int tns_plan_stage_export(const tns_plan* p, int stage, int64_t* records);

#ifdef __cplusplus
}
#endif

#endif /* TNSGPU_H */
