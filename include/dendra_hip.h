/*
 * dendra_hip.h -- C ABI of libdnhip.so, the gfx950 cable hot path.
 *
 * Conventions for every entry point:
 *  - pointers are device pointers (hipMalloc / torch CUDA tensors) unless
 *    marked "host"; sizes are element counts;
 *  - inputs are never modified; outputs are caller-allocated;
 *  - work is enqueued asynchronously on `stream` (0 = null stream); no entry
 *    point synchronises the device or allocates device memory;
 *  - return 0 on success, DN_EINVAL (-1) for invalid arguments, DN_ELAUNCH
 *    (-2) when a kernel launch fails, DN_ERANGE (-3) when a size exceeds what
 *    the kernel supports without the optional caller scratch.
 * Citations name the reference interface each entry point replaces
 * (paths relative to the wmglab-duke/dendra repository root).
 */
#ifndef DENDRA_HIP_H
#define DENDRA_HIP_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DN_OK 0
#define DN_EINVAL -1
#define DN_ELAUNCH -2
#define DN_ERANGE -3
#define DN_EUNSUPPORTED -4

/* ---- batched tridiagonal solve ------------------------------------------
 * Replaces thomas_solve_cuda_t / ThomasSolve._solve
 * (dendra/models/integrators/triton/t_kernel_thomas.py:83-130,218-232) and
 * pcr_solve_cuda_t (triton/pcr_kernel_thomas.py:99-195,224-226).
 * a, c: (B, K-1); b, d, x: (B, K), row-major.  method: 0 auto, 1 Thomas
 * (lane per system; K <= 32 in float32 or 16 in float64 stages the
 * coefficients in LDS and needs no scratch), 2 wavefront hybrid PCR.  The
 * adjoint solve of
 * ThomasSolve.backward (t_kernel_thomas.py:147-189) is the same call with a
 * and c exchanged.  `scratch` holds dn_tridiag_scratch_elems() values (may
 * be NULL when that is 0). */
int64_t dn_tridiag_scratch_elems(int64_t B, int64_t K, int method, int elem_bytes);
int dn_tridiag_solve_f32(const float* a, const float* b, const float* c,
                         const float* d, float* x, int64_t B, int64_t K, int method,
                         float* scratch, int64_t scratch_elems, hipStream_t stream);
int dn_tridiag_solve_f64(const double* a, const double* b, const double* c,
                         const double* d, double* x, int64_t B, int64_t K, int method,
                         double* scratch, int64_t scratch_elems, hipStream_t stream);

/* ---- tree (DHS / Hines) solve ---------------------------------------------
 * Replaces dhs_solve_cuda / _single_dhs_kernel(_packed)
 * (dendra/models/integrators/triton/dhs_kernel.py:14-212,493-651), bound by
 * _dhs._select_solver (integrators/tree.py:812-839).  d_mem, a_geom, b, x:
 * (B, K) in solver order; parent (K), order (K), layer_ptr (L+1) shared by
 * the batch, int64, as built by build_dhs_layers (tree.py:615-673). */
int64_t dn_tree_dhs_scratch_elems(int64_t B, int64_t K, int elem_bytes);
int dn_tree_dhs_f32(const float* d_mem, const float* a_geom, const float* b,
                    const int64_t* parent, const int64_t* order,
                    const int64_t* layer_ptr, float* x, int64_t B, int64_t K, int64_t L,
                    float* scratch, int64_t scratch_elems, hipStream_t stream);
int dn_tree_dhs_f64(const double* d_mem, const double* a_geom, const double* b,
                    const int64_t* parent, const int64_t* order,
                    const int64_t* layer_ptr, double* x, int64_t B, int64_t K, int64_t L,
                    double* scratch, int64_t scratch_elems, hipStream_t stream);

/* ---- tree solve by contraction --------------------------------------------
 * Replaces the tree-contraction solve(d, g, rhs, plan)
 * (dendra/models/integrators/triton/tree_contraction.py:748-770; the
 * integrator factory dn.tree_contraction, integrators/__init__.py:73-110,
 * binds it for _dhs._step, integrators/tree.py:186-189,841-867).  d, g, rhs,
 * x: (B, K) in the caller's node order; g is the edge conductance to the
 * parent (to ground at the root).  The plan tables (section chains and the
 * rake/compress rounds over them) are int32 device arrays built by
 * dendra_amd.contraction.TreeContractionPlan: perm (K), sec_* (n_sec),
 * elim_ptr (n_rounds + 1) over elim_sec/elim_par/elim_child, recv_ptr
 * (n_rounds + 1) over recv_sec, src_ptr (n_recv + 1) over src.  work holds
 * dn_tree_contract_work_elems(B, K) elements of the solve's dtype.  Returns
 * DN_ERANGE when n_sec exceeds dn_tree_contract_max_sections. */
int64_t dn_tree_contract_work_elems(int64_t B, int64_t K);
int dn_tree_contract_max_sections(int elem_bytes);
int dn_tree_contract_f32(const float* d, const float* g, const float* rhs, float* x,
                         int64_t B, int64_t K, const int32_t* perm, int n_sec,
                         const int32_t* sec_head, const int32_t* sec_len,
                         const int32_t* sec_child, const int32_t* sec_nchild,
                         const int32_t* sec_parent, int n_rounds, const int32_t* elim_ptr,
                         const int32_t* elim_sec, const int32_t* elim_par,
                         const int32_t* elim_child, const int32_t* recv_ptr,
                         const int32_t* recv_sec, const int32_t* src_ptr,
                         const int32_t* src, float* work, int64_t work_elems,
                         hipStream_t stream);
int dn_tree_contract_f64(const double* d, const double* g, const double* rhs, double* x,
                         int64_t B, int64_t K, const int32_t* perm, int n_sec,
                         const int32_t* sec_head, const int32_t* sec_len,
                         const int32_t* sec_child, const int32_t* sec_nchild,
                         const int32_t* sec_parent, int n_rounds, const int32_t* elim_ptr,
                         const int32_t* elim_sec, const int32_t* elim_par,
                         const int32_t* elim_child, const int32_t* recv_ptr,
                         const int32_t* recv_sec, const int32_t* src_ptr,
                         const int32_t* src, double* work, int64_t work_elems,
                         hipStream_t stream);

/* ---- 3x3 block solvers (double-layer extracellular model) -----------------
 * Replace the block solvers that _bwd_euler_bt (integrators/implicit.py:
 * 669-1136) and the block tree integrator (integrators/tree_bt.py) select:
 *   dn_bt3_thomas_*  thomas_bt3_solve / thomas_solve_cuda_bt
 *                    (triton/bt_kernel.py:192-298): main (B,K,3,3), lower and
 *                    upper (B,K-1,3) diagonal coupling blocks, rhs/x (B,K,3);
 *   dn_bt3_spd_*     solve_bt_spd_fwd_consume_impl (factor = 1: factors main,
 *                    writes chol (B,K,9) and per-system info = blocks that
 *                    stayed indefinite after the diagonal shifts) and
 *                    solve_bt_spd_solve_with_chol_impl (factor = 0: reuses
 *                    chol; main may be null) (triton/bt_spd_kernel.py:575-700);
 *   dn_bt3_tree_*    _dhs_bt3_adjoint (triton/dhs_kernel_bt.py): main
 *                    (B,K,3,3), edge (B,K,3) per-component conductance to the
 *                    parent, parent (K), order (K, children before parents),
 *                    layer_ptr (n_layers + 1) chunking `order` into runs of
 *                    one depth, and the children CSR child_ptr (K + 1) /
 *                    child_idx; with the layers and CSR given (and layers
 *                    averaging two or more nodes) one wavefront solves a
 *                    system layer by layer, else (layer_ptr null) one lane
 *                    walks `order`.  When every layer is one whole depth
 *                    level of at most 64 nodes (checked on the device), the
 *                    children's terms pass through LDS instead of `work`
 *                    (same solution to rounding); passing whole depth levels
 *                    therefore pays.
 * Inputs are not modified; work holds dn_block_work_elems(kind, B, K)
 * elements (kind 0 Thomas, 1 SPD, 2 tree). */
int64_t dn_block_work_elems(int kind, int64_t B, int64_t K);
int dn_bt3_thomas_f32(const float* lower, const float* main, const float* upper,
                      const float* rhs, float* x, int64_t B, int64_t K, float* work,
                      int64_t work_elems, hipStream_t stream);
int dn_bt3_thomas_f64(const double* lower, const double* main, const double* upper,
                      const double* rhs, double* x, int64_t B, int64_t K, double* work,
                      int64_t work_elems, hipStream_t stream);
int dn_bt3_spd_f32(const float* lower, const float* main, const float* upper,
                   const float* rhs, float* x, float* chol, int32_t* info, int64_t B,
                   int64_t K, float* work, int64_t work_elems, int factor,
                   hipStream_t stream);
int dn_bt3_spd_f64(const double* lower, const double* main, const double* upper,
                   const double* rhs, double* x, double* chol, int32_t* info, int64_t B,
                   int64_t K, double* work, int64_t work_elems, int factor,
                   hipStream_t stream);
int dn_bt3_tree_f32(const float* main, const float* edge, const float* rhs,
                    const int64_t* parent, const int64_t* order, const int64_t* layer_ptr,
                    int64_t n_layers, const int32_t* child_ptr, const int32_t* child_idx,
                    float* x, int64_t B, int64_t K, float* work, int64_t work_elems,
                    hipStream_t stream);
int dn_bt3_tree_f64(const double* main, const double* edge, const double* rhs,
                    const int64_t* parent, const int64_t* order, const int64_t* layer_ptr,
                    int64_t n_layers, const int32_t* child_ptr, const int32_t* child_idx,
                    double* x, int64_t B, int64_t K, double* work, int64_t work_elems,
                    hipStream_t stream);

/* dn_bt3_tree_contract_*: the same block tree system (gmem, nullable: a
 * membrane conductance g per node added to its main block as
 * g [[1,-1,0],[-1,1,0],[0,0,0]]) solved by contraction over a section plan,
 * the 3x3-block form of dn_tree_contract_* (same tables: perm, n_sec
 * sections sec_head/sec_len/sec_child/sec_nchild/sec_parent, n_rounds
 * rake/compress rounds: elim_ptr .. src; see dendra_amd.contraction):
 * each section eliminates its nodes head first, then the section tree is
 * contracted, then everything is recovered.  Main blocks need not be
 * symmetric.  One workgroup per system; DN_ERANGE when n_sec exceeds
 * dn_bt3_contract_max_sections(element bytes, K); work holds
 * dn_bt3_contract_work_elems(B, K) elements. */
int dn_bt3_contract_max_sections(int elem_bytes, int64_t K);
int64_t dn_bt3_contract_work_elems(int64_t B, int64_t K);
int dn_bt3_tree_contract_f32(const float* main, const float* edge, const float* rhs,
                             const float* gmem, float* x, int64_t B, int64_t K,
                             const int32_t* perm, int n_sec, const int32_t* sec_head,
                             const int32_t* sec_len, const int32_t* sec_child,
                             const int32_t* sec_nchild, const int32_t* sec_parent, int n_rounds,
                             const int32_t* elim_ptr, const int32_t* elim_sec,
                             const int32_t* elim_par, const int32_t* elim_child,
                             const int32_t* recv_ptr, const int32_t* recv_sec,
                             const int32_t* src_ptr, const int32_t* src, float* work,
                             int64_t work_elems, hipStream_t stream);
int dn_bt3_tree_contract_f64(const double* main, const double* edge, const double* rhs,
                             const double* gmem, double* x, int64_t B, int64_t K,
                             const int32_t* perm, int n_sec, const int32_t* sec_head,
                             const int32_t* sec_len, const int32_t* sec_child,
                             const int32_t* sec_nchild, const int32_t* sec_parent, int n_rounds,
                             const int32_t* elim_ptr, const int32_t* elim_sec,
                             const int32_t* elim_par, const int32_t* elim_child,
                             const int32_t* recv_ptr, const int32_t* recv_sec,
                             const int32_t* src_ptr, const int32_t* src, double* work,
                             int64_t work_elems, hipStream_t stream);

/* ---- NetCon delivery with fractional (interpolated) delays ----------------
 * Replaces the ordered two-tap enqueue (dendra/models/networks/
 * _two_tap_enqueue_cpu.cpp:26-152, dendra_native::ordered_two_tap_enqueue)
 * and the three-tap scatter of dense_interpolated_delay_transition
 * (networks/_event_delivery.py:226-310).  q (depth, width) receives the
 * transition of qin (qin null or equal to q: in place; else qin is read and
 * left unchanged): row *slot (device int64, reduced modulo depth) is
 * cleared; then
 *   taps 1: every connection adds w g at row slot + base (f unused);
 *   taps 2: every connection adds w g (1 - f) at row slot + base and w g f at
 *           row slot + base + 1, all lower taps before all upper taps;
 *   taps 3: every connection adds w g f[3e + j] at row slot + base + j,
 *           j = 0..2, connection-major;
 * rows modulo depth, each cell in ascending connection order (bitwise the
 * reference's sequence).  in_ptr (width + 1) / in_edge (n_edge) list each
 * target's connections in ascending order; w, g, base (n_edge) and f
 * (n_edge, or 3 n_edge) are per connection; base is the first tap's row
 * offset modulo depth.  dn_interp_delay_vjp_* writes the per-connection
 * cotangents dw, dg, df for a ring cotangent G.
 * dn_interp_delay_enqueue_ws_*: the same transition with a workspace of
 * dn_interp_delay_work_bytes(n_edge, taps, f64) bytes (n_edge = number of
 * connections; in_edge must then be a permutation of [0, n_edge), i.e.
 * in_ptr[width] == n_edge): each connection's tap values and first row are
 * packed into one record, read in connection order and written at the
 * connection's CSR position pos[e] (pos[in_edge[p]] = p; pass it once
 * computed, e.g. cached with the CSR, or NULL to have it computed in the
 * workspace each call), so the target-major pass reads the records in order
 * instead of gathering w, g, f and base; results are identical bit for
 * bit. */
int dn_interp_delay_enqueue_f32(const float* qin, float* q, int64_t depth, int64_t width,
                                const int64_t* slot, int taps, const int32_t* in_ptr,
                                const int32_t* in_edge, const float* w, const float* g,
                                const float* f, const int32_t* base, hipStream_t stream);
int dn_interp_delay_enqueue_f64(const double* qin, double* q, int64_t depth, int64_t width,
                                const int64_t* slot, int taps, const int32_t* in_ptr,
                                const int32_t* in_edge, const double* w, const double* g,
                                const double* f, const int32_t* base, hipStream_t stream);
int64_t dn_interp_delay_work_bytes(int64_t n_edge, int taps, int f64);
int dn_interp_delay_enqueue_ws_f32(const float* qin, float* q, int64_t depth, int64_t width,
                                   const int64_t* slot, int taps, const int32_t* in_ptr,
                                   const int32_t* in_edge, const int32_t* pos, const float* w,
                                   const float* g,
                                   const float* f, const int32_t* base, int64_t n_edge,
                                   void* work, int64_t work_bytes, hipStream_t stream);
int dn_interp_delay_enqueue_ws_f64(const double* qin, double* q, int64_t depth, int64_t width,
                                   const int64_t* slot, int taps, const int32_t* in_ptr,
                                   const int32_t* in_edge, const int32_t* pos, const double* w,
                                   const double* g,
                                   const double* f, const int32_t* base, int64_t n_edge,
                                   void* work, int64_t work_bytes, hipStream_t stream);
int dn_interp_delay_vjp_f32(const float* G, int64_t depth, int64_t width,
                            const int64_t* slot, int taps, int64_t n_edge,
                            const int32_t* post, const float* w, const float* g,
                            const float* f, const int32_t* base, float* dw, float* dg,
                            float* df, hipStream_t stream);
int dn_interp_delay_vjp_f64(const double* G, int64_t depth, int64_t width,
                            const int64_t* slot, int taps, int64_t n_edge,
                            const int32_t* post, const double* w, const double* g,
                            const double* f, const int32_t* base, double* dw, double* dg,
                            double* df, hipStream_t stream);

/* ---- source-history transport ---------------------------------------------
 * Replaces source_history_project / source_history_record
 * (dendra/models/networks/_source_history.py:83-186), the transport of the
 * source-history training backend (networks/netcon.py:3660-3725).
 *   dn_history_project_*: hist (depth, n_col) floating gates (packed = 0,
 *     element type of w) or int64 words (packed = 1, gate = (word & mask[e])
 *     != 0); connection e reads column col[e] at rows (slot - off[e][j]) mod
 *     depth for j < taps (off int32 in [0, depth), taps 1 fixed, 2 linear
 *     with upper weight frac[e], 3 Gaussian with frac[e][3]) and
 *     out[t] = sum over the target-major CSR (in_ptr, in_edge) of
 *     w[e] * mixture, accumulated from 0 in ascending connection order.
 *   dn_history_project_ws_*: the same projection with a workspace of
 *     dn_history_work_bytes(n_conn, taps, packed, f64, depth, n_col) bytes
 *     (n_conn = number of connections; in_edge a permutation of
 *     [0, n_conn); n_col = 0 gives the size without the copy below): each
 *     connection's column, tap rows, tap weights, w (and mask) are packed
 *     into one record written at the connection's CSR position pos[e]
 *     (pos[in_edge[p]] = p, passed in or NULL to compute it in the
 *     workspace), so the target-major pass reads the records in order; a
 *     floating history is first copied column-major into the workspace
 *     (when work_bytes leaves room for it, as the size returned does), so
 *     a connection's adjacent tap rows share a cache line; results are
 *     identical bit for bit.  The records hold tap offsets, not rows: with
 *     reuse != 0 the call reads the records an earlier call left in work
 *     (same in_edge, pos, col, mask, off, w, frac and sizes; hist and slot
 *     may differ), as the steps of a simulation with fixed weights and
 *     delays do; reuse = 0 packs them.
 *   dn_history_project_vjp_*: per connection dw, dfrac (taps 2, 3) and, for
 *     a floating history, dhist += (atomic; zero it first, may be null).
 *   dn_history_record_*: row slot mod depth := gates (in place).
 *   dn_history_record_packed: row slot mod depth := sum over sources with
 *     active[s] != 0 of mask[s] at word[s] (in place, int64 sums).
 * slot is a one-element device int64. */
int dn_history_project_f32(const void* hist, int packed, int64_t depth, int64_t n_col,
                           const int64_t* slot, int taps, const int32_t* in_ptr,
                           const int32_t* in_edge, const int32_t* col, const int64_t* mask,
                           const int32_t* off, const float* w, const float* frac,
                           float* out, int64_t width, hipStream_t stream);
int dn_history_project_f64(const void* hist, int packed, int64_t depth, int64_t n_col,
                           const int64_t* slot, int taps, const int32_t* in_ptr,
                           const int32_t* in_edge, const int32_t* col, const int64_t* mask,
                           const int32_t* off, const double* w, const double* frac,
                           double* out, int64_t width, hipStream_t stream);
int64_t dn_history_work_bytes(int64_t n_conn, int taps, int packed, int f64, int64_t depth,
                              int64_t n_col);
int dn_history_project_ws_f32(const void* hist, int packed, int64_t depth, int64_t n_col,
                              const int64_t* slot, int taps, const int32_t* in_ptr,
                              const int32_t* in_edge, const int32_t* pos, const int32_t* col,
                              const int64_t* mask,
                              const int32_t* off, const float* w, const float* frac,
                              float* out, int64_t width, int64_t n_conn, void* work,
                              int64_t work_bytes, int reuse,
                              hipStream_t stream);
int dn_history_project_ws_f64(const void* hist, int packed, int64_t depth, int64_t n_col,
                              const int64_t* slot, int taps, const int32_t* in_ptr,
                              const int32_t* in_edge, const int32_t* pos, const int32_t* col,
                              const int64_t* mask,
                              const int32_t* off, const double* w, const double* frac,
                              double* out, int64_t width, int64_t n_conn, void* work,
                              int64_t work_bytes, int reuse,
                              hipStream_t stream);
int dn_history_project_vjp_f32(const float* grad_out, const void* hist, int packed,
                               int64_t depth, int64_t n_col, const int64_t* slot, int taps,
                               int64_t n_conn, const int32_t* post, const int32_t* col,
                               const int64_t* mask, const int32_t* off, const float* w,
                               const float* frac, float* dw, float* dfrac, float* dhist,
                               hipStream_t stream);
int dn_history_project_vjp_f64(const double* grad_out, const void* hist, int packed,
                               int64_t depth, int64_t n_col, const int64_t* slot, int taps,
                               int64_t n_conn, const int32_t* post, const int32_t* col,
                               const int64_t* mask, const int32_t* off, const double* w,
                               const double* frac, double* dw, double* dfrac, double* dhist,
                               hipStream_t stream);
int dn_history_record_f32(float* hist, int64_t depth, int64_t n_col, const int64_t* slot,
                          const float* gates, hipStream_t stream);
int dn_history_record_f64(double* hist, int64_t depth, int64_t n_col, const int64_t* slot,
                          const double* gates, hipStream_t stream);
int dn_history_record_packed(int64_t* hist, int64_t depth, int64_t n_words,
                             const int64_t* slot, int64_t n_src, const uint8_t* active,
                             const int32_t* word, const int64_t* mask, hipStream_t stream);

/* ---- pending calendar of the sparse NetCon backend ------------------------
 * Replaces pending_calendar_append / pending_calendar_pop
 * (dendra/models/networks/_pending_calendar.py:37-113), the storage of the
 * sparse_calendar delay backend (networks/netcon.py:3309-3450, 5223-5300).
 * A calendar is slots/indices (int64) and values (float, double or int32) of
 * `cap` entries, a device int64 size and a device bool overflow flag.
 *   dn_calendar_append_*: slots/indices/values hold a copy of the old
 *     calendar and receive, in order, the candidates with active[i] != 0 at
 *     old_size + rank -- unless old_ovf or old_size + count > cap, when
 *     nothing is written; new_size/new_ovf receive the result; dest (may be
 *     null) the position of each candidate or -1.
 *   dn_calendar_pop_*: entries p < size with slots[p] == *slot are due;
 *     payload[indices[p]] += values[p] in position order (payload holds its
 *     initial value; indices outside [0, width) are dropped); the other valid
 *     entries go, in order, to the front of n_*, whose remaining entries are
 *     zeroed;
 *     new_size = their count; keys_out (may be null) gets index for due
 *     entries and width otherwise, keep_dest (may be null) each entry's new
 *     position or -1.
 * work: dn_calendar_work_bytes(n, cap, kind) bytes (kind 0 f32, 1 f64,
 * 2 i32; n = candidates, 0 for a pop).  A pop given
 * dn_calendar_pop_work_bytes(cap, width, kind) bytes adds the due values
 * through per-index hit lists (each due entry files its position at its
 * index; one thread per hit index adds them in position order) instead of
 * a stable radix sort of all cap keys; the payload is the same bit for bit. */
int64_t dn_calendar_work_bytes(int64_t n, int64_t cap, int kind);
int64_t dn_calendar_pop_work_bytes(int64_t cap, int64_t width, int kind);
int dn_calendar_append_f32(int64_t* slots, int64_t* indices, float* values, int64_t cap,
                          const int64_t* old_size, const uint8_t* old_ovf,
                          int64_t* new_size, uint8_t* new_ovf, int64_t n,
                          const int64_t* c_slots, const int64_t* c_indices,
                          const float* c_values, const uint8_t* active, int64_t* dest,
                          void* work, int64_t work_bytes, hipStream_t stream);
int dn_calendar_pop_f32(const int64_t* slots, const int64_t* indices, const float* values,
                       int64_t cap, const int64_t* size, const int64_t* slot, int64_t width,
                       int64_t* n_slots, int64_t* n_indices, float* n_values,
                       int64_t* new_size, float* payload, int32_t* keys_out,
                       int64_t* keep_dest, void* work, int64_t work_bytes,
                       hipStream_t stream);
int dn_calendar_append_f64(int64_t* slots, int64_t* indices, double* values, int64_t cap,
                          const int64_t* old_size, const uint8_t* old_ovf,
                          int64_t* new_size, uint8_t* new_ovf, int64_t n,
                          const int64_t* c_slots, const int64_t* c_indices,
                          const double* c_values, const uint8_t* active, int64_t* dest,
                          void* work, int64_t work_bytes, hipStream_t stream);
int dn_calendar_pop_f64(const int64_t* slots, const int64_t* indices, const double* values,
                       int64_t cap, const int64_t* size, const int64_t* slot, int64_t width,
                       int64_t* n_slots, int64_t* n_indices, double* n_values,
                       int64_t* new_size, double* payload, int32_t* keys_out,
                       int64_t* keep_dest, void* work, int64_t work_bytes,
                       hipStream_t stream);
int dn_calendar_append_i32(int64_t* slots, int64_t* indices, int32_t* values, int64_t cap,
                          const int64_t* old_size, const uint8_t* old_ovf,
                          int64_t* new_size, uint8_t* new_ovf, int64_t n,
                          const int64_t* c_slots, const int64_t* c_indices,
                          const int32_t* c_values, const uint8_t* active, int64_t* dest,
                          void* work, int64_t work_bytes, hipStream_t stream);
int dn_calendar_pop_i32(const int64_t* slots, const int64_t* indices, const int32_t* values,
                       int64_t cap, const int64_t* size, const int64_t* slot, int64_t width,
                       int64_t* n_slots, int64_t* n_indices, int32_t* n_values,
                       int64_t* new_size, int32_t* payload, int32_t* keys_out,
                       int64_t* keep_dest, void* work, int64_t work_bytes,
                       hipStream_t stream);

/* ---- HH cables with two extracellular layers (ExtCell) --------------------
 * Replaces _bwd_euler_bt._step (dendra/models/integrators/implicit.py:
 * 1021-1136) with its block solve (thomas_solve_cuda_bt) for B HH cables of
 * K compartments, all n_steps in one launch.  Layout [compartment][component]
 * [cell]: maind (K,9,B), lower/upper (K-1,3,B) (diagonal couplings at
 * (k,k-1) / (k-1,k)), area (K,B), xg radial conductances (K,2,B), optional
 * sna/sk/lead/iamp (K,B), amp (B), wave/i_wave (n_steps); state vc (K,3,B)
 * = (vi, ve0, ve1) and m/h/n (K,B) advance in place; rec (optional)
 * receives vi - ve0 every rec_every steps, (n_rec,K,B).  work:
 * dn_extcell_work_elems(B, K) elements. */
int64_t dn_extcell_work_elems(int64_t B, int64_t K);
int dn_extcell_run_f32(int64_t B, int64_t K, const float* maind, const float* lower,
                       const float* upper, const float* area, const float* xg,
                       const float* sna, const float* sk, const float* hh, const float* lead,
                       const float* amp, const float* wave, const float* iamp,
                       const float* i_wave, float dt, int64_t n_steps, float* vc, float* m,
                       float* h, float* n, float* rec, int64_t rec_every, float* work,
                       int64_t work_elems, hipStream_t stream);
int dn_extcell_run_f64(int64_t B, int64_t K, const double* maind, const double* lower,
                       const double* upper, const double* area, const double* xg,
                       const double* sna, const double* sk, const double* hh,
                       const double* lead, const double* amp, const double* wave,
                       const double* iamp, const double* i_wave, double dt, int64_t n_steps,
                       double* vc, double* m, double* h, double* n, double* rec,
                       int64_t rec_every, double* work, int64_t work_elems,
                       hipStream_t stream);

/* ---- HH trees with two extracellular layers (ExtCell trees) ---------------
 * Replaces _dhs_bt._step (dendra/models/integrators/tree_bt.py:504-654) with
 * its block tree solve (_dhs_bt3_adjoint) for B trees of one topology, all
 * n_steps in one call, arrays in node order: main (B,K,9) radial main blocks
 * and edge (B,K,3) conductances to the parent (tree_bt.py:288-372), area
 * (B,K), xg radial conductances (B,K,2), parent32/parent (K, int32/int64,
 * -1 at a root), order and layer_ptr (n_layers + 1) the DHS layers, child_ptr
 * (K + 1) / child_idx the children CSR (ascending child index); optional
 * sna/sk/lead/iamp (B,K), amp (B), wave/i_wave (n_steps); state vc (B,K,3),
 * m/h/n (B,K) in place; rec (optional) vi - ve0 every rec_every steps,
 * (n_rec,B,K).  work: dn_extcell_tree_work_elems(B, K) elements.
 * dn_extcell_tree_run_contract_*: the same run with the block tree solve by
 * contraction over a section plan (the tables of dn_bt3_tree_contract_*) in
 * place of parent/order/layer_ptr; its work holds
 * dn_extcell_tree_contract_work_elems(B, K) elements. */
int64_t dn_extcell_tree_work_elems(int64_t B, int64_t K);
int64_t dn_extcell_tree_contract_work_elems(int64_t B, int64_t K);
int dn_extcell_tree_run_f32(int64_t B, int64_t K, const float* main, const float* edge,
                            const float* area, const float* xg, const int32_t* parent32,
                            const int64_t* parent, const int64_t* order,
                            const int64_t* layer_ptr, int64_t n_layers,
                            const int32_t* child_ptr, const int32_t* child_idx,
                            const float* sna, const float* sk, const float* hh, const float* lead,
                            const float* amp, const float* wave, const float* iamp,
                            const float* i_wave, float dt, int64_t n_steps, float* vc, float* m,
                            float* h, float* n, float* rec, int64_t rec_every, float* work,
                            int64_t work_elems, hipStream_t stream);
int dn_extcell_tree_run_f64(int64_t B, int64_t K, const double* main, const double* edge,
                            const double* area, const double* xg, const int32_t* parent32,
                            const int64_t* parent, const int64_t* order,
                            const int64_t* layer_ptr, int64_t n_layers,
                            const int32_t* child_ptr, const int32_t* child_idx,
                            const double* sna, const double* sk, const double* hh, const double* lead,
                            const double* amp, const double* wave, const double* iamp,
                            const double* i_wave, double dt, int64_t n_steps, double* vc, double* m,
                            double* h, double* n, double* rec, int64_t rec_every, double* work,
                            int64_t work_elems, hipStream_t stream);
int dn_extcell_tree_run_contract_f32(
    int64_t B, int64_t K, const float* main, const float* edge, const float* area,
    const float* xg, const int32_t* parent32, const int32_t* child_ptr, const int32_t* child_idx,
    const int32_t* perm, int n_sec, const int32_t* sec_head, const int32_t* sec_len,
    const int32_t* sec_child, const int32_t* sec_nchild, const int32_t* sec_parent, int n_rounds,
    const int32_t* elim_ptr, const int32_t* elim_sec, const int32_t* elim_par,
    const int32_t* elim_child, const int32_t* recv_ptr, const int32_t* recv_sec,
    const int32_t* src_ptr, const int32_t* src, const float* sna, const float* sk,
    const float* hh, const float* lead, const float* amp, const float* wave, const float* iamp,
    const float* i_wave, float dt, int64_t n_steps, float* vc, float* m, float* h, float* n,
    float* rec, int64_t rec_every, float* work, int64_t work_elems, hipStream_t stream);
int dn_extcell_tree_run_contract_f64(
    int64_t B, int64_t K, const double* main, const double* edge, const double* area,
    const double* xg, const int32_t* parent32, const int32_t* child_ptr,
    const int32_t* child_idx, const int32_t* perm, int n_sec, const int32_t* sec_head,
    const int32_t* sec_len, const int32_t* sec_child, const int32_t* sec_nchild,
    const int32_t* sec_parent, int n_rounds, const int32_t* elim_ptr, const int32_t* elim_sec,
    const int32_t* elim_par, const int32_t* elim_child, const int32_t* recv_ptr,
    const int32_t* recv_sec, const int32_t* src_ptr, const int32_t* src, const double* sna,
    const double* sk, const double* hh, const double* lead, const double* amp,
    const double* wave, const double* iamp, const double* i_wave, double dt, int64_t n_steps,
    double* vc, double* m, double* h, double* n, double* rec, int64_t rec_every, double* work,
    int64_t work_elems, hipStream_t stream);

/* ---- fused HH + backward Euler, unbranched cables -------------------------
 * Replaces the per-step transition of _bwd_euler_ub
 * (dendra/models/integrators/implicit.py:561-664) in Integrator._step order
 * (integrators/core.py:769-834) with HH gating (mod/HH.py:24-116,
 * mechanisms/_cnexp.py:96-132), driven from Population.run
 * (models/core.py:4555-4614).  Ragged batch in one flat compartment axis,
 * cells in plan order with offsets `off` (n_cells+1).  Per compartment:
 * gL = g_j/C_j, gRp = g_{j-1}/C_j (0 at cable ends), scale = area/C,
 * drive = activating-function coefficient scaled by wave[step] (nullable),
 * sna/sk gNa/gK scales (nullable).  hh (host): {gnabar, gkbar, gl, ena, ek,
 * el, q10}.  dt in ms.  rec (nullable): V after every rec_every-th absolute
 * step, row (step0+s+1)/rec_every-1 of a [rows][total] array.  max_rows
 * (1..dn_hh_cable_rmax()) bounds task_nr over the launch and selects the
 * kernel instance (8, 10 or 13 rows per lane); a task with more rows than
 * the instance holds is left unadvanced and sets bit 1 of *status.
 * status (nullable): caller-owned device int32 that the kernel ORs launch-
 * contract violations into (1 cable rows, 2 tree group over threads or node
 * capacity, 4 tree over kmax/smax/lmax); the caller reads it after the stream
 * has run (dendra_amd.population does, and raises DN_ERANGE's RuntimeError). */
int dn_hh_cable_rmax(void);
int dn_hh_cable_run_f32(int n_tasks, const int32_t* task_cell0,
                        const int32_t* task_ncell, const int32_t* task_nr,
                        const int32_t* task_p, const int64_t* off, float* v, float* m,
                        float* h, float* n, const float* gL, const float* gRp,
                        const float* scale, const float* drive, const float* sna,
                        const float* sk, const float* wave, int64_t step0,
                        int64_t n_steps, float dt, const float* hh, float* rec,
                        int64_t rec_every, int64_t total, int waves_per_block,
                        int max_rows, int32_t* status, hipStream_t stream);
int dn_hh_cable_run_f64(int n_tasks, const int32_t* task_cell0,
                        const int32_t* task_ncell, const int32_t* task_nr,
                        const int32_t* task_p, const int64_t* off, double* v, double* m,
                        double* h, double* n, const double* gL, const double* gRp,
                        const double* scale, const double* drive, const double* sna,
                        const double* sk, const double* wave, int64_t step0,
                        int64_t n_steps, double dt, const double* hh, double* rec,
                        int64_t rec_every, int64_t total, int waves_per_block,
                        int max_rows, int32_t* status, hipStream_t stream);
/* One plain step (gating+assembly launch, solve launch) of
 * _bwd_euler_ub._step (integrators/core.py:769-834, implicit.py:569-664);
 * mainv/rhs are caller scratch of `total` values; w is this step's waveform
 * value for `drive`.  ve (nullable): this step's extracellular potential
 * per compartment (mV), entering through its axial differences; intra
 * (nullable): injected current per compartment (mA), scaled by icap = 1/C
 * (1/F; required with intra).  gx, gex (nullable, together): further affine
 * membrane terms per compartment, G += gx (S/cm^2), J += gex (S/cm^2 mV),
 * e.g. exp2syn conductances from dn_exp2syn_membrane_*.  v is advanced in
 * place. */
int dn_hh_cable_step_f32(int64_t n_cells, const int64_t* off, int64_t total, int max_nr,
                         float* v, float* m, float* h, float* n, const float* gL,
                         const float* gRp, const float* scale, const float* drive,
                         const float* sna, const float* sk, float w, float dt,
                         const float* hh, const float* ve, const float* intra,
                         const float* icap, const float* gx, const float* gex, float* mainv,
                         float* rhs, hipStream_t stream);
int dn_hh_cable_step_f64(int64_t n_cells, const int64_t* off, int64_t total, int max_nr,
                         double* v, double* m, double* h, double* n, const double* gL,
                         const double* gRp, const double* scale, const double* drive,
                         const double* sna, const double* sk, double w, double dt,
                         const double* hh, const double* ve, const double* intra,
                         const double* icap, const double* gx, const double* gex, double* mainv,
                         double* rhs,
                         hipStream_t stream);

/* ---- fused HH + backward Euler, branched trees ----------------------------
 * Replaces the per-step transition of _dhs (integrators/tree.py:1202-1357)
 * and the packed multi-tree path _dhs_multi / dhs_solve_multi_cuda
 * (tree.py:2097-2334, triton/dhs_kernel_multi.py:16-135,831-888).  Trees in
 * section-chain numbering (sections of <= 16 nodes, contiguous, breadth-first
 * over the section tree): parent/ch_ptr/ch_idx and the section arrays
 * (head, len, first child section, child count, parent node of the head) are
 * local to each tree; lvl_ptr holds per-tree section offsets per level;
 * smax / lmax bound every tree's section and level counts.  Per node:
 * area, cmdt = C/dt_s, g = child-indexed axial conductance (0 at the root),
 * drive (mA per unit waveform, nullable).  The persistent run launches the
 * `n_launch` trees listed in `order` with one block of `threads` threads per
 * tree; layout 1 (one thread per section, state in registers) takes kmax =
 * the launch's largest section count <= threads <= 512, layout 0 (eight node
 * slots per thread) requires
 * kmax >= each tree's size and kmax <= 8 * threads.  Work outside those
 * bounds is left unadvanced and reported in *status (nullable device int32,
 * bits as for dn_hh_cable_run_*). */
int dn_hh_tree_slots(void);
/* Longest section (unbranched chain) the tree kernels take; the host's
 * section plan must split longer chains.  Layout-1 runs report longer
 * sections through *status (bit 2) and leave their group unadvanced. */
int dn_hh_tree_sec_max(void);
int dn_hh_tree_run_f32(int n_trees, const int64_t* node_off, const int32_t* parent,
                       const int32_t* ch_ptr, const int32_t* ch_idx,
                       const int32_t* sec_off, const int32_t* sec_head,
                       const int32_t* sec_len, const int32_t* sec_child,
                       const int32_t* sec_nchild, const int32_t* sec_ptail,
                       const int32_t* sec_parent, const int32_t* lvl_off,
                       const int32_t* lvl_ptr, int smax, int lmax,
                       const int32_t* order, int n_launch, const int32_t* grp_off,
                       int n_groups, int layout,
                       int kmax, int threads, float* v, float* m, float* h, float* n,
                       const float* area, const float* cmdt, const float* g,
                       const float* drive, const float* sna, const float* sk,
                       const float* wave, int64_t step0, int64_t n_steps, float dt,
                       const float* hh, float* rec, int64_t rec_every, int64_t total,
                       int32_t* status, hipStream_t stream);
int dn_hh_tree_run_f64(int n_trees, const int64_t* node_off, const int32_t* parent,
                       const int32_t* ch_ptr, const int32_t* ch_idx,
                       const int32_t* sec_off, const int32_t* sec_head,
                       const int32_t* sec_len, const int32_t* sec_child,
                       const int32_t* sec_nchild, const int32_t* sec_ptail,
                       const int32_t* sec_parent, const int32_t* lvl_off,
                       const int32_t* lvl_ptr, int smax, int lmax,
                       const int32_t* order, int n_launch, const int32_t* grp_off,
                       int n_groups, int layout,
                       int kmax, int threads, double* v, double* m, double* h, double* n,
                       const double* area, const double* cmdt, const double* g,
                       const double* drive, const double* sna, const double* sk,
                       const double* wave, int64_t step0, int64_t n_steps, double dt,
                       const double* hh, double* rec, int64_t rec_every, int64_t total,
                       int32_t* status, hipStream_t stream);
/* One plain step of _dhs._step (integrators/tree.py:1202-1357): per-node
 * gating and the current-form residual (mA), then the section-chain solve.
 * ve (nullable): this step's extracellular potential per node (mV), as edge
 * currents g (ve_child - ve_parent); intra (nullable): injected current per
 * node (mA); gx, gex (nullable, together): affine membrane terms per node as
 * for dn_hh_cable_step_*.  v is advanced in place. */
int dn_hh_tree_step_f32(int n_trees, const int64_t* node_off, const int32_t* parent,
                        const int32_t* ch_ptr, const int32_t* ch_idx,
                        const int32_t* sec_off, const int32_t* sec_head,
                        const int32_t* sec_len, const int32_t* sec_child,
                        const int32_t* sec_nchild, const int32_t* sec_ptail,
                        const int32_t* sec_parent, const int32_t* lvl_off,
                        const int32_t* lvl_ptr, int smax, int lmax,
                        const int32_t* node_tree, int kmax,
                        int64_t total, float* v, float* m, float* h, float* n,
                        const float* area, const float* cmdt, const float* g,
                        const float* drive, const float* sna, const float* sk, float w,
                        float dt, const float* hh, const float* ve, const float* intra,
                        const float* gx, const float* gex, float* mainv, float* rhs,
                        hipStream_t stream);
int dn_hh_tree_step_f64(int n_trees, const int64_t* node_off, const int32_t* parent,
                        const int32_t* ch_ptr, const int32_t* ch_idx,
                        const int32_t* sec_off, const int32_t* sec_head,
                        const int32_t* sec_len, const int32_t* sec_child,
                        const int32_t* sec_nchild, const int32_t* sec_ptail,
                        const int32_t* sec_parent, const int32_t* lvl_off,
                        const int32_t* lvl_ptr, int smax, int lmax,
                        const int32_t* node_tree, int kmax,
                        int64_t total, double* v, double* m, double* h, double* n,
                        const double* area, const double* cmdt, const double* g,
                        const double* drive, const double* sna, const double* sk,
                        double w, double dt, const double* hh, const double* ve,
                        const double* intra, const double* gx, const double* gex,
                        double* mainv, double* rhs, hipStream_t stream);

/* ---- event delivery ---------------------------------------------------------
 * Delivery from a spike tape.  Replaces the bitpacked source history and its
 * delivery rebuild, pack_source_spikes / build_delivery(_uniform)
 * (dendra/models/networks/netcon_bitpack_kernel.cu:20-126,
 * netcon_bitpack_kernel.cpp:32-149, called from netcon.py:3739-3883), and the
 * dense ring of NetCon.advance_non_diff_dense_* (netcon.py:5183-5205,
 * _event_delivery.py:30-199).  Connections are numbered target-major
 * (in_ptr, n_targets+1); out_ptr/out_tgt/out_e is the by-source CSR (target
 * and connection index of each entry).  Per step the ring row `ctr[0]` is
 * copied to delivered[step][*] and cleared, then the connections of the
 * sources set in tape[step][*] add w into ring[(cur + delay) % depth][post]
 * in connection order (bitwise equal to the ordered float32 scatter_add).
 * Scratch: spk_bits ceil(n_src/32) words, spk_list n_src, hit_cnt n_targets
 * (zero on entry; zero again after every step), hit_e 16 * n_targets
 * (16-byte aligned),
 * tgt_list n_targets, ctr {cur, step, 0, 0}. */
int dn_net_deliver_tape_f32(int n_targets, int n_src, const int32_t* out_ptr,
                            const int32_t* out_tgt, const int32_t* out_e,
                            const int32_t* in_ptr, const int32_t* pre, const int32_t* delay,
                            const float* w, int depth, float* ring, uint32_t* spk_bits,
                            int32_t* spk_list, int32_t* hit_cnt, int32_t* hit_e,
                            int32_t* tgt_list, int32_t* ctr, const uint8_t* tape,
                            float* delivered, int64_t n_steps, hipStream_t stream);

/* ---- network step ------------------------------------------------------------
 * Replaces the Network.run hot loop (dendra/models/networks/net.py:4485-4545,
 * 335-377) over dense NetCons, exp2syn receivers (mod/EXP2SYN.py:112-156) and
 * HH populations stepped by _bwd_euler_ub / _bwd_euler_sc
 * (integrators/implicit.py:121-200,569-664).  Cells [0, n_point) have one
 * compartment, the rest up to 256; one exp2syn per cell at flat index
 * soma[c]; sources are the cells (rising edge at theta) then Poisson sources
 * with p = bg_p per step (splitmix64 counter RNG of seed, bg_stream).
 * syn (host) = {factor, tau1_eff, tau2, e}.  Optional records: ev_log
 * (step, source) pairs up to ev_cap, delivered payloads for the first
 * rec_steps steps.  Delivery arrays and scratch as dn_net_deliver_tape_f32.
 * ctr: 8 int32 {cur, step, 0, ...}.  graph_steps > 0
 * captures that many steps as one HIP graph and replays it (stream must not
 * be the null stream). */
int dn_net_run_f32(int n_cells, int n_point, const int64_t* off, const int32_t* soma,
                   float* v, float* m, float* h, float* n, const float* gL,
                   const float* gRp, const float* scale, const float* sna,
                   const float* sk, const float* area, float dt, const float* hh,
                   float* A, float* B, const float* syn, int n_src, uint8_t* latch,
                   float theta, uint64_t seed, uint32_t bg_stream, double bg_p,
                   const int32_t* out_ptr, const int32_t* out_tgt, const int32_t* out_e,
                   const int32_t* in_ptr, const int32_t* pre, const int32_t* delay,
                   const float* w, int depth, float* ring, uint32_t* spk_bits,
                   int32_t* spk_list, int32_t* hit_cnt, int32_t* hit_e, int32_t* tgt_list,
                   int32_t* ctr, int32_t* ev_log, int32_t* ev_count,
                   int64_t ev_cap, float* delivered, int64_t rec_steps, int64_t n_steps,
                   int graph_steps, hipStream_t stream);

/* ---- trace-fitting loss and its hand-written adjoint ---------------------
 * Replace autograd through Population.run (dendra/models/core.py:4555-4614)
 * for HH cables driven by a piecewise-linear Ve knot table (waveform
 * `arbitrary`, dendra/models/stim/waveform/implementations.py:748-796):
 * per step ThomasSolve.backward's transposed solve
 * (dendra/models/integrators/triton/t_kernel_thomas.py:147-189) plus the
 * gating vector-Jacobian products autograd would record
 * (mechanisms/_cnexp.py:96-132, mod/HH.py:24-56,109-116), with checkpointed
 * recomputation as Population.longrun_checkpointed (core.py:7802).
 * Slot layout: task t is one workgroup of `waves` wavefronts (W = 64*waves
 * threads) and owns W*rcap slots; slot t*W*rcap + r*W + tid is row r of
 * thread tid.  (rcap, waves) is (4, 1) or (8, 1) -- a wavefront of cells of
 * task_p in {4,8,16,32,64} lanes -- or (2, 2) or (1, 4): one cell over the
 * W threads (task_p = W).  Per-slot
 * arrays are [slots], state/adjoint [4][slots] (v, m, h, n), knots
 * [n_knots][slots], trajectory [n_steps+1][4][slots].  act holds the
 * activating function of each Ve knot (dn_fit_csr_knots_*).  knot_k/knot_a
 * (length >= step0+n_steps) give the segment and weight of step n; target
 * is [steps][tstride], site[] its column for recorded rows (-1 otherwise),
 * mw[] the margin weight.  hh (host) = {gnabar, gkbar, gl, ena, ek, el, q10}.
 * loss (optional, [n_tasks*waves][2]) receives per-wavefront (mse, margin)
 * sums. */
int dn_fit_forward_f32(int n_tasks, int rcap, int waves, const int32_t* task_nr,
                       const int32_t* task_p,
                       const float* gl, const float* grp, const float* scale,
                       const float* sna, const float* sk, const float* act, int n_knots,
                       const int32_t* knot_k, const float* knot_a, const int32_t* site,
                       const float* mw, const float* target, int64_t tstride,
                       float mse_scale, float theta, float tau, int64_t step0,
                       int64_t n_steps, float dt, const float* hh, const float* x0,
                       float* xout, float* ckpt, int64_t ckpt_every, float* traj,
                       float* loss, hipStream_t stream);
int dn_fit_forward_f64(int n_tasks, int rcap, int waves, const int32_t* task_nr,
                       const int32_t* task_p,
                       const double* gl, const double* grp, const double* scale,
                       const double* sna, const double* sk, const double* act,
                       int n_knots, const int32_t* knot_k, const double* knot_a,
                       const int32_t* site, const double* mw, const double* target,
                       int64_t tstride, double mse_scale, double theta, double tau,
                       int64_t step0, int64_t n_steps, double dt, const double* hh,
                       const double* x0, double* xout, double* ckpt, int64_t ckpt_every,
                       double* traj, double* loss, hipStream_t stream);
/* Reverse sweep over steps [step0, step0+n_steps) from the trajectory the
 * forward wrote for that segment; adj carries (dL/dv, dL/dm, dL/dh, dL/dn)
 * in and out; g_sna, g_sk, g_act are accumulated (+=). */
int dn_fit_adjoint_f32(int n_tasks, int rcap, int waves, const int32_t* task_nr,
                       const int32_t* task_p,
                       const float* gl, const float* grp, const float* scale,
                       const float* sna, const float* sk, const float* act, int n_knots,
                       const int32_t* knot_k, const float* knot_a, const int32_t* site,
                       const float* mw, const float* target, int64_t tstride,
                       float mse_scale, float theta, float tau, int64_t step0,
                       int64_t n_steps, float dt, const float* hh, const float* traj,
                       float* adj, float* g_sna, float* g_sk, float* g_act,
                       hipStream_t stream);
int dn_fit_adjoint_f64(int n_tasks, int rcap, int waves, const int32_t* task_nr,
                       const int32_t* task_p,
                       const double* gl, const double* grp, const double* scale,
                       const double* sna, const double* sk, const double* act,
                       int n_knots, const int32_t* knot_k, const double* knot_a,
                       const int32_t* site, const double* mw, const double* target,
                       int64_t tstride, double mse_scale, double theta, double tau,
                       int64_t step0, int64_t n_steps, double dt, const double* hh,
                       const double* traj, double* adj, double* g_sna, double* g_sk,
                       double* g_act, hipStream_t stream);
/* out[i*out_row + k*out_k] = sum_{e in [ptr[i], ptr[i+1])} w[e] *
 * in[idx[e]*in_row + k*in_k] for k < n_k: the activating-function operator
 * (implicit.py:618-638) applied to every Ve knot, and its transpose for the
 * knot gradients, in a fixed summation order (deterministic). */
int dn_fit_csr_knots_f32(int64_t n_out, int n_k, const int64_t* ptr, const int32_t* idx,
                         const float* w, const float* in, int64_t in_row, int64_t in_k,
                         float* out, int64_t out_row, int64_t out_k, hipStream_t stream);
int dn_fit_csr_knots_f64(int64_t n_out, int n_k, const int64_t* ptr, const int32_t* idx,
                         const double* w, const double* in, int64_t in_row, int64_t in_k,
                         double* out, int64_t out_row, int64_t out_k, hipStream_t stream);

/* ---- activation thresholds by bisection ----------------------------------
 * Replaces Thresholder.calculate_thresholds with _fix_bounds (block_possible
 * False) and arithmetic calc_stimamp (dendra/models/instruments/
 * thresholder.py:835-993) under the Active callback (models/callbacks.py:
 * 1946-1990) for unbranched HH cables with Ve_j(t) = lead_j * amp * wave[t].
 * Slot layout with rcap 4 (slot task*256 + r*64 + lane); dr holds the
 * activating function of the unit lead field, chk flags checked rows.
 * ub/lb (plan order): initial bounds in, final bounds out; status 0
 * converged, 1 bound fixing failed, 2 tolerance not met; trials = trials run.
 * hh (host) = {gnabar, gkbar, gl, ena, ek, el, q10}; state0 (host) = rest
 * {v, m, h, n}.  Each wavefront runs all trials of its fibers in one launch. */
int dn_threshold_f64(int n_tasks, int rcap, const int32_t* task_nr, const int32_t* task_p,
                     const int32_t* task_cell0, const int32_t* task_ncell,
                     const double* gl, const double* grp, const double* scale,
                     const double* dr, const int32_t* chk, const double* wave,
                     int64_t n_steps, double dt, const double* hh, const double* state0,
                     double vthr, double rtol, double fix_up, int max_fix, int max_bisect,
                     double* ub, double* lb, int32_t* status, int32_t* trials,
                     hipStream_t stream);

/* ---- threshold surrogate (BF16 MFMA multilayer perceptron) ---------------
 * No counterpart in the reference: the learned part that predicts the
 * thresholds dn_threshold_f64 computes, used with bisection as fallback.
 * BF16 operands are passed as void* (2-byte elements); strides are in
 * elements; z indexes ensemble members.  dn_mlp_gemm: C = A B^T for nz
 * products, A [M][K] and B [N][K] BF16 rows, float32 accumulation on
 * v_mfma_f32_32x32x16_bf16; epi 0: Z = C + bias, H = gelu(Z) (BF16), HT =
 * H^T; epi 1: H = C * gelu'(Zp), HT = H^T; epi 2: split-K float32 partials
 * Z[z*splits + s].  M, N multiples of 128, K a multiple of 32*splits. */
int dn_mlp_gemm(int epi, int nz, int M, int N, int K, int splits, const void* A, int64_t lda,
                int64_t sa, const void* B, int64_t ldb, int64_t sb, const float* bias,
                int64_t sbias, float* Z, int64_t ldz, int64_t sz, void* H, int64_t ldh,
                int64_t sh, void* HT, int64_t ldht, int64_t sht, const float* Zp,
                int64_t ldzp, int64_t szp, hipStream_t stream);
/* Output unit y = H3 . w4 + b4; with targets t also dy = 2 (y - t) / Bn. */
int dn_mlp_out(int nz, int Bn, int Hd, const void* H3, int64_t sh, const void* w4, int64_t sw,
               const float* b4, const float* t, int64_t st, float* y, float* dy, int64_t sy,
               hipStream_t stream);
/* dZ3 = dy w4^T * gelu'(Z3) in [b][k] and [k][b] BF16 layouts. */
int dn_mlp_out_bwd(int nz, int Bn, int Hd, const float* dy, int64_t sy, const void* w4,
                   int64_t sw, const float* Z3, int64_t sz, void* dZ, void* dZT, int64_t sd,
                   hipStream_t stream);
/* out[z][r] = sum_c X[z][r][c] * (w ? w[z][c] : 1) (bias and output-weight
 * gradients), X BF16 when x_bf16 else float32. */
int dn_mlp_row_dot(int nz, int rows, int cols, int x_bf16, const void* X, int64_t sx,
                   const float* w, int64_t sw, float* out, int64_t so, hipStream_t stream);
/* out[z][i] = sum_s P[z][s][i] over split-K partials. */
int dn_mlp_splitk_reduce(int nz, int64_t n, int splits, const float* P, float* out,
                         hipStream_t stream);
/* One Adam step on float32 masters with bias-corrected step lr_t; refreshes
 * the BF16 copy pb, the transposed BF16 copy pbt, and the copies in MFMA
 * fragment order of W (pbp, rows % 32 == 0, cols % 16 == 0; read by
 * dn_mlp_forward) and of W^T (pbtp, cols % 32 == 0, rows % 16 == 0; read by
 * dn_mlp_backward); any of the four may be NULL. */
int dn_mlp_adam(int64_t n, int rows, int cols, float* p, const float* g, float* m, float* v,
                float lr_t, float b1, float b2, float eps, void* pb, void* pbt, void* pbp,
                void* pbtp, hipStream_t stream);
/* dn_mlp_adam over `count` <= 8 tensors in one launch: per tensor k the
 * arguments of dn_mlp_adam as host arrays (n, rows, cols; device addresses
 * p, g, m, v and the copies pb, pbt, pbp, pbtp, 0 for none). */
int dn_mlp_adam_multi(int count, const int64_t* n, const int* rows, const int* cols,
                      const uint64_t* p, const uint64_t* g, const uint64_t* m,
                      const uint64_t* v, const uint64_t* pb, const uint64_t* pbt,
                      const uint64_t* pbp, const uint64_t* pbtp, float lr_t, float b1,
                      float b2, float eps, hipStream_t stream);
/* The whole forward in one launch per 128-row tile and member: x[128] ->
 * 512 -> 512 -> 512 (GELU) -> 1, activations kept in LDS, weights Wp1..Wp3
 * in MFMA fragment order ([z][N/32][K/16][64][8]: element e of lane l in
 * block (nb, ks) is W[32 nb + l%32][16 ks + 8 (l/32) + e]; member strides
 * sw1, sw23 in elements).  X [z][n][128] (sx = 0: one X for every member),
 * n a multiple of 128; y [z][n] (stride sy).  Inference: Z*, HT*, t, dy
 * NULL; tile = 128 or 64 rows per workgroup.  Training (tile 64): also, in
 * Z1..Z3 [z][n][512] float32, gelu'(Z_l) of the pre-activations (the only
 * use the backward makes of them), HT_l [z][512][n] BF16 activations and
 * dy = 2 (y - t) / n, as the layer products dn_mlp_gemm (epi 0) +
 * dn_mlp_out produce them, and into part [z][n/64][5][512] the tile sums
 * of dw4 = H3^T dy (slot 3) and db4 = sum dy (slot 4, element 0). */
int dn_mlp_forward(int nz, int n, int tile, const void* X, int64_t sx, const void* Wp1,
                   const void* Wp2, const void* Wp3, int64_t sw1, int64_t sw23,
                   const float* b1, const float* b2, const float* b3, const void* w4,
                   const float* b4, float* y, int64_t sy, float* Z1, float* Z2, float* Z3,
                   void* HT1, void* HT2, void* HT3, const float* t, float* dy, float* part,
                   hipStream_t stream);
/* The backward data products of a training step in one launch per 64-row
 * tile (tile must be 64), from the gelu'(Z_l) that dn_mlp_forward leaves in
 * Z1..Z3: dZ3 = dy w4^T * gelu'(Z3), dZ2 = (dZ3 W3) *
 * gelu'(Z2), dZ1 = (dZ2 W2) * gelu'(Z1) in BF16 as dn_mlp_out_bwd and
 * dn_mlp_gemm (epi 1) produce them, each kept in LDS for the next product;
 * W2^T, W3^T packed in MFMA fragment order (member stride sw).  Writes
 * dZ_l^T [z][512][n] and the tile column sums of dZ_l into part slots 0..2. */
int dn_mlp_backward(int nz, int n, int tile, const float* dy, int64_t sy, const void* w4,
                    const float* Z1, const float* Z2, const float* Z3, const void* Wt2p,
                    const void* Wt3p, int64_t sw, void* dZ1T, void* dZ2T, void* dZ3T,
                    float* part, hipStream_t stream);
/* Gradients of b1, b2, b3, w4 ([z][512]) and b4 ([z]) from the tile
 * partials [z][tiles][5][512], summed over the tiles in order. */
int dn_mlp_partials_reduce(int nz, int tiles, const float* part, float* gb1, float* gb2,
                           float* gb3, float* gw4, float* gb4, hipStream_t stream);
/* Features x[0..P-1] = profile / max|profile|, x[P] = (diam - dc) / ds, zero
 * padding to ldx; lmag = log max|profile|; with lo/hi (P+2 bounds, the last
 * for lmag) flag = 1 where a value is outside [lo - margin, hi + margin]. */
int dn_mlp_featurize(int64_t n, int P, const float* prof, const float* diam, float dc, float ds,
                     void* X, int ldx, float* lmag, const float* lo, const float* hi,
                     float margin, int32_t* flag, hipStream_t stream);
/* Minibatch gather X0[z][b] = X[idx[z][b]] in both layouts, t0 = t[idx]. */
int dn_mlp_gather(int nz, int Bn, int F, const int32_t* idx, const void* X, const float* t,
                  void* X0, void* X0T, float* t0, hipStream_t stream);
/* thr = median over members of exp(pred * ystd + ymean - lmag); flag = 1
 * where (max - min) > spread * median (flags already set stay set). */
int dn_mlp_ensemble(int64_t n, int nz, const float* pred, int64_t sp, const float* lmag,
                    float ymean, float ystd, float spread, float* thr, int32_t* flag,
                    hipStream_t stream);

/* ---- packed source history and dense delivery (NetCon bitpack) -----------
 * Replaces the reference's netcon_bitpack extension: pack_source_spikes
 * (networks/netcon_bitpack_kernel.cu:20-57,130-148; .cpp:32-48) and
 * build_delivery / build_delivery_uniform (.cu:59-126,150-228; .cpp:50-143),
 * called through networks/netcon_bitpack_ops.py:559-700 by
 * netcon.py:3758-3827.  history: int64 [depth][n_words], n_words =
 * ceil(n_source / 63); word w holds sources 63w..63w+62 in bits 0..62.
 * cur: device int64[1], the ring row of this step.
 * dn_bitpack_pack writes row cur from spikes (bool bytes, n_source).
 * dn_bitpack_deliver_* adds weight[e] into out[post[e]] for every connection
 * whose bit (word_idx[e], bit_mask[e]) is set in row (cur - delay[e]) mod
 * depth; delay NULL means one delay for all, uniform_delay (below 1 -> 1, at
 * or beyond depth -> modulo depth, as the reference).  out (n_out values) is
 * added to, not cleared; hits on one target add in an unspecified order.
 * Words, targets or rows out of range are skipped. */
int dn_bitpack_pack(const uint8_t* spikes, int64_t n_source, int64_t* history,
                    int64_t depth, int64_t n_words, const int64_t* cur, hipStream_t stream);
int dn_bitpack_deliver_f32(const int64_t* history, int64_t depth, int64_t n_words,
                           const int64_t* cur, const int64_t* delay, int64_t uniform_delay,
                           const int64_t* word_idx, const int64_t* bit_mask,
                           const int64_t* post, const float* weight, float* out,
                           int64_t n_out, int64_t n_conn, hipStream_t stream);
int dn_bitpack_deliver_f64(const int64_t* history, int64_t depth, int64_t n_words,
                           const int64_t* cur, const int64_t* delay, int64_t uniform_delay,
                           const int64_t* word_idx, const int64_t* bit_mask,
                           const int64_t* post, const double* weight, double* out,
                           int64_t n_out, int64_t n_conn, hipStream_t stream);

/* ---- dense delay ring with fixed integer delays ---------------------------
 * The fixed-delay transition dense_fixed_delay_transition
 * (networks/_event_delivery.py:125-199, the default dense backend of
 * NetCon.advance_non_diff_dense_uniform/_mixed, netcon.py:5183-5205) is
 * dn_interp_delay_enqueue_* with taps = 1, base[e] = delay[e] mod depth and
 * f = NULL: every ring cell receives w[e] * g[e] in connection order; its
 * cotangent is dn_interp_delay_vjp_* with taps = 1 (f, df NULL).
 * dn_dense_event_step: the int32 event-count ring [depth][n_conn] of the same
 * transition (:175-190): qout = qin with row slot cleared, then
 * qout[(slot + delay[e]) mod depth][e] += counts[e].
 * dn_ordered_scatter_add_*: _enqueue_dense_values (:61-76), q[dest[i]] +=
 * values[i] for i in index order per cell (bitwise the sequential
 * scatter_add); dest outside [0, n_cells) is dropped; `work` holds
 * dn_ordered_scatter_work_bytes(n, n_cells, f64) bytes. */
int dn_dense_event_step(const int32_t* qin, int32_t* qout, int64_t depth, int64_t n_conn,
                        const int64_t* slot, const int64_t* delay, const int32_t* counts,
                        hipStream_t stream);
int64_t dn_ordered_scatter_work_bytes(int64_t n, int64_t n_cells, int f64);
int dn_ordered_scatter_add_f32(float* q, int64_t n_cells, const int64_t* dest,
                               const float* values, int64_t n, void* work, int64_t work_bytes,
                               hipStream_t stream);
int dn_ordered_scatter_add_f64(double* q, int64_t n_cells, const int64_t* dest,
                               const double* values, int64_t n, void* work,
                               int64_t work_bytes, hipStream_t stream);

/* ---- exp2syn point processes, per step ------------------------------------
 * Replaces the reference's exp2syn mechanism (mod/EXP2SYN.py) inside
 * Network.step (networks/net.py:335-377) for callers that keep that order:
 * dn_exp2syn_receive_*: net_receive_values (:154-156, Synapse.net_receive,
 * mechanisms/_mechanism.py:4518-4524), A += payload * factor, B likewise
 * (factor per synapse, or factor_all when NULL);
 * dn_exp2syn_advance_*: the cnexp step of A' = -A/tau1, B' = -B/tau2
 * (mechanisms/_cnexp.py:96-132), taus per synapse or the *_all scalars;
 * dn_exp2syn_membrane_*: the affine current (B - A)(v - e) nA (:147-152) as
 * membrane densities for the step's gx, gex: for each listed compartment
 * comp[c] (CSR comp_ptr over syn_order, synapse order within), g = sum
 * (B - A) 1e-6 / area[comp] and g e; other compartments are not written. */
int dn_exp2syn_receive_f32(float* A, float* B, const float* payload, const float* factor,
                           float factor_all, int64_t n, hipStream_t stream);
int dn_exp2syn_receive_f64(double* A, double* B, const double* payload,
                           const double* factor, double factor_all, int64_t n,
                           hipStream_t stream);
int dn_exp2syn_advance_f32(float* A, float* B, const float* tau1, float tau1_all,
                           const float* tau2, float tau2_all, float dt, int64_t n,
                           hipStream_t stream);
int dn_exp2syn_advance_f64(double* A, double* B, const double* tau1, double tau1_all,
                           const double* tau2, double tau2_all, double dt, int64_t n,
                           hipStream_t stream);
int dn_exp2syn_membrane_f32(const float* A, const float* B, const float* e, float e_all,
                            const float* area, const int32_t* comp_ptr, const int32_t* comp,
                            const int32_t* syn_order, int64_t n_comp, float* g_out,
                            float* ge_out, hipStream_t stream);
int dn_exp2syn_membrane_f64(const double* A, const double* B, const double* e,
                            double e_all, const double* area, const int32_t* comp_ptr,
                            const int32_t* comp, const int32_t* syn_order, int64_t n_comp,
                            double* g_out, double* ge_out, hipStream_t stream);

/* ---- packed multi-morphology tree solve ----------------------------------
 * Replaces dhs_solve_multi_cuda / DHSSolveMultiPacked and
 * _multi_dhs_grad_a_kernel (integrators/triton/dhs_kernel_multi.py:16-206,
 * 629-888; caller _dhs_multi._step, integrators/tree.py:2284-2300).  d, a, b,
 * x, D_work: (B_total, K_stride) row-padded planes in solver order.  Topology
 * tables (int64): P_cat / order_cat (n_cat) and lptr_cat (n_lcat), local to
 * each morphology; plan entry w (int64 tables of n_warps) covers rows
 * warp_row_base[w] .. + warp_row_count[w] - 1 of the morphology at offsets
 * warp_p_off / warp_order_off / warp_lptr_off with warp_l layers (at most
 * L_max visited).  x must be zero on entry: entries no plan row covers stay
 * zero.  D_work: caller workspace of the same shape (overwritten).  max_rows
 * sizes the workgroup (rows per entry; larger entries loop).  grad_a: the
 * gradient with respect to a for the solution V and the adjoint solution G,
 * written on covered entries (dA zero on entry). */
int dn_tree_dhs_multi_f32(const float* d, const float* a, const float* b,
                          const int64_t* P_cat, const int64_t* order_cat, int64_t n_cat,
                          const int64_t* lptr_cat, int64_t n_lcat, const int64_t* warp_p_off,
                          const int64_t* warp_order_off, const int64_t* warp_lptr_off,
                          const int64_t* warp_l, const int64_t* warp_row_base,
                          const int64_t* warp_row_count, int64_t n_warps, int64_t B_total,
                          int64_t K_stride, int64_t L_max, int max_rows, float* D_work,
                          float* x, hipStream_t stream);
int dn_tree_dhs_multi_f64(const double* d, const double* a, const double* b,
                          const int64_t* P_cat, const int64_t* order_cat, int64_t n_cat,
                          const int64_t* lptr_cat, int64_t n_lcat, const int64_t* warp_p_off,
                          const int64_t* warp_order_off, const int64_t* warp_lptr_off,
                          const int64_t* warp_l, const int64_t* warp_row_base,
                          const int64_t* warp_row_count, int64_t n_warps, int64_t B_total,
                          int64_t K_stride, int64_t L_max, int max_rows, double* D_work,
                          double* x, hipStream_t stream);
int dn_tree_dhs_multi_grad_a_f32(const float* V, const float* G, const int64_t* P_cat,
                                 const int64_t* order_cat, int64_t n_cat,
                                 const int64_t* lptr_cat, int64_t n_lcat,
                                 const int64_t* warp_p_off, const int64_t* warp_order_off,
                                 const int64_t* warp_lptr_off, const int64_t* warp_l,
                                 const int64_t* warp_row_base, const int64_t* warp_row_count,
                                 int64_t n_warps, int64_t B_total, int64_t K_stride,
                                 int64_t L_max, float* dA, hipStream_t stream);
int dn_tree_dhs_multi_grad_a_f64(const double* V, const double* G, const int64_t* P_cat,
                                 const int64_t* order_cat, int64_t n_cat,
                                 const int64_t* lptr_cat, int64_t n_lcat,
                                 const int64_t* warp_p_off, const int64_t* warp_order_off,
                                 const int64_t* warp_lptr_off, const int64_t* warp_l,
                                 const int64_t* warp_row_base, const int64_t* warp_row_count,
                                 int64_t n_warps, int64_t B_total, int64_t K_stride,
                                 int64_t L_max, double* dA, hipStream_t stream);

/* ---- allocation-free tridiagonal solve (float32) --------------------------
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same solve and arguments as dn_tridiag_solve_f32
 * without the caller scratch: the workspace is a library-owned device array
 * of dn_tridiag_noalloc_max_scratch_elems() floats (8 MiB) per GPU, held
 * until the process ends.  Returns DN_ERANGE when the solve would need more
 * (dn_tridiag_scratch_elems(B, K, method, 4) above that size).  Calls no
 * allocating routine.  Calls that overlap on one GPU share the array and race. */
int64_t dn_tridiag_noalloc_max_scratch_elems(void);
int dn_tridiag_solve_noalloc_f32(const float* a, const float* b, const float* c,
                                 const float* d, float* x, int64_t B, int64_t K, int method,
                                 hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* DENDRA_HIP_H */
