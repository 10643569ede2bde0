// dn_syn.hip -- exp2syn point processes as a per-step receiver.
//
// GPU form of the reference's exp2syn mechanism (dendra/models/mod/EXP2SYN.py)
// for callers that keep Dendra's Network.step order (networks/net.py:335-377:
// NetCon delivery into the synapses, then the populations' step):
//   receive    net_receive_values (:154-156): A += w factor, B += w factor for
//              the payload w delivered to each synapse (Synapse.net_receive,
//              mechanisms/_mechanism.py:4518-4524);
//   advance    the cnexp step of A' = -A / tau1_effective, B' = -B / tau2
//              (mechanisms/_cnexp.py:96-132: the Taylor form for dt/tau <= 0.5,
//              the exact exponential beyond), run where the integrator
//              advances its mechanisms before the membrane frame;
//   membrane   the affine current i = (B - A)(v - e) nA (:147-152) as
//              per-compartment density terms for the membrane frame:
//              g = sum (B - A) 1e-6 / area (S/cm^2) and g e, summed over each
//              compartment's synapses in synapse order (a compartment CSR).
// One lane per synapse (receive, advance) or per compartment (membrane).
#include "dn_common.hpp"

namespace dn {

template <typename T>
__global__ __launch_bounds__(256) void exp2syn_receive_kernel(
    T* __restrict__ A, T* __restrict__ B, const T* __restrict__ payload,
    const T* __restrict__ factor, T factor_all, int64_t n) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T w = payload[i] * (factor ? factor[i] : factor_all);
  A[i] = A[i] + w;
  B[i] = B[i] + w;
}

template <typename T>
__global__ __launch_bounds__(256) void exp2syn_advance_kernel(
    T* __restrict__ A, T* __restrict__ B, const T* __restrict__ tau1, T tau1_all,
    const T* __restrict__ tau2, T tau2_all, T dt, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T k1 = T(1) / (tau1 ? tau1[i] : tau1_all);
  const T k2 = T(1) / (tau2 ? tau2[i] : tau2_all);
  A[i] = cnexp(A[i], T(0), k1, dt);
  B[i] = cnexp(B[i], T(0), k2, dt);
}

template <typename T>
__global__ __launch_bounds__(256) void exp2syn_membrane_kernel(
    const T* __restrict__ A, const T* __restrict__ B, const T* __restrict__ e, T e_all,
    const T* __restrict__ area, const int32_t* __restrict__ comp_ptr,
    const int32_t* __restrict__ comp, const int32_t* __restrict__ syn_order,
    int64_t n_comp, T* __restrict__ g_out, T* __restrict__ ge_out) {
#pragma clang fp contract(off)
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_comp) return;
  const int32_t k = comp[c];
  const T scale = T(1e-6) / area[k];
  T g = T(0), ge = T(0);
  for (int32_t q = comp_ptr[c]; q < comp_ptr[c + 1]; ++q) {
    const int32_t s = syn_order[q];
    const T gs = (B[s] - A[s]) * scale;
    g = g + gs;
    ge = ge + gs * (e ? e[s] : e_all);
  }
  g_out[k] = g;
  ge_out[k] = ge;
}

}  // namespace dn

namespace {

template <typename T>
int receive_impl(T* A, T* B, const T* payload, const T* factor, T factor_all, int64_t n,
                 hipStream_t stream) {
  if (n < 0) return DN_EINVAL;
  if (n == 0) return DN_OK;
  if (!A || !B || !payload) return DN_EINVAL;
  const int64_t blocks = (n + 255) / 256;
  if (blocks > INT32_MAX) return DN_ERANGE;
  hipLaunchKernelGGL((dn::exp2syn_receive_kernel<T>), dim3((unsigned)blocks), dim3(256), 0,
                     stream, A, B, payload, factor, factor_all, n);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

template <typename T>
int advance_impl(T* A, T* B, const T* tau1, T tau1_all, const T* tau2, T tau2_all, T dt,
                 int64_t n, hipStream_t stream) {
  if (n < 0) return DN_EINVAL;
  if (n == 0) return DN_OK;
  if (!A || !B) return DN_EINVAL;
  if ((!tau1 && !(tau1_all > T(0))) || (!tau2 && !(tau2_all > T(0)))) return DN_EINVAL;
  const int64_t blocks = (n + 255) / 256;
  if (blocks > INT32_MAX) return DN_ERANGE;
  hipLaunchKernelGGL((dn::exp2syn_advance_kernel<T>), dim3((unsigned)blocks), dim3(256), 0,
                     stream, A, B, tau1, tau1_all, tau2, tau2_all, dt, n);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

template <typename T>
int membrane_impl(const T* A, const T* B, const T* e, T e_all, const T* area,
                  const int32_t* comp_ptr, const int32_t* comp, const int32_t* syn_order,
                  int64_t n_comp, T* g_out, T* ge_out, hipStream_t stream) {
  if (n_comp < 0) return DN_EINVAL;
  if (n_comp == 0) return DN_OK;
  if (!A || !B || !area || !comp_ptr || !comp || !syn_order || !g_out || !ge_out)
    return DN_EINVAL;
  const int64_t blocks = (n_comp + 255) / 256;
  if (blocks > INT32_MAX) return DN_ERANGE;
  hipLaunchKernelGGL((dn::exp2syn_membrane_kernel<T>), dim3((unsigned)blocks), dim3(256), 0,
                     stream, A, B, e, e_all, area, comp_ptr, comp, syn_order, n_comp, g_out,
                     ge_out);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

}  // namespace

#define DN_SYN_ENTRIES(SFX, T)                                                              \
  DN_EXPORT int dn_exp2syn_receive_##SFX(T* A, T* B, const T* payload, const T* factor,     \
                                         T factor_all, int64_t n, hipStream_t stream) {     \
    return receive_impl<T>(A, B, payload, factor, factor_all, n, stream);                   \
  }                                                                                         \
  DN_EXPORT int dn_exp2syn_advance_##SFX(T* A, T* B, const T* tau1, T tau1_all,             \
                                         const T* tau2, T tau2_all, T dt, int64_t n,        \
                                         hipStream_t stream) {                              \
    return advance_impl<T>(A, B, tau1, tau1_all, tau2, tau2_all, dt, n, stream);            \
  }                                                                                         \
  DN_EXPORT int dn_exp2syn_membrane_##SFX(const T* A, const T* B, const T* e, T e_all,      \
                                          const T* area, const int32_t* comp_ptr,           \
                                          const int32_t* comp, const int32_t* syn_order,    \
                                          int64_t n_comp, T* g_out, T* ge_out,              \
                                          hipStream_t stream) {                             \
    return membrane_impl<T>(A, B, e, e_all, area, comp_ptr, comp, syn_order, n_comp, g_out, \
                            ge_out, stream);                                                \
  }

DN_SYN_ENTRIES(f32, float)
DN_SYN_ENTRIES(f64, double)
