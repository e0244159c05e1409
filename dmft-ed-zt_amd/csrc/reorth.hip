// FP64 reorthogonalization of a Lanczos chain against its stored Krylov
// vectors (an option of the Green's-function chains; the reference runs the
// plain three-term recurrence, lanczos_solver.f90:147-170).
//
// After step j's axpy the basis slot j+1 holds r = H q_j - alpha_j q_j -
// beta_j q_{j-1} (unnormalized u_{j+1}), with q_k = s_k u_k in slots 0..j.
// Three kernels follow the axpy:
//   k_ro_dot     every workgroup decides whether this step reorthogonalizes:
//                always (full mode), or, in partial mode, when Simon's
//                omega recurrence (H. D. Simon, Math. Comp. 42 (1984) 115)
//                estimates some |<q_{j+1}, q_k>| above sqrt(eps), and at the
//                step after such a step.  Workgroup 0 stores the new omega
//                row and the decision.  If it reorthogonalizes, each
//                workgroup writes FP64 partial dots <u_k, r> of its chunk.
//   k_ro_reduce  c_k = s_k^2 times the sum of the partials (fixed order).
//   k_ro_update  r <- r - sum_k c_k u_k in FP64, rounded to the storage
//                type, and the partial sums of |r|^2 that the next step's
//                H.v prologue reduces into beta_{j+1} (the axpy's partials
//                stay when the step does not reorthogonalize).
// Classical Gram-Schmidt in one pass: the stored vectors are orthogonal to
// the storage precision when it runs, and omega is reset to eps after it.
#include "device_util.h"

namespace edg {

constexpr int kRoBlock = 256;
constexpr int kRoUpdBlock = 1024;
constexpr int kRoGroup = 8;  // stored vectors per pass of k_ro_dot and k_ro_update

// Omega row j+1 from rows j and j-1 (Simon 1984, the omega recurrence),
// for k < j:
//   beta_{j+1} w'_k = beta_{k+1} w_{k+1} + (alpha_k - alpha_j) w_k + beta_k w_{k-1} - beta_j w''_k
// plus a rounding term sign * eps * |T|, w'_j = eps |T| / beta_{j+1}, w'_{j+1} = 1.
template <typename T>
__global__ __launch_bounds__(kRoBlock) void k_ro_dot(ReorthArgs a, const T *r, const T *base, long long stride,
                                                     long long n)
{
    if (*a.stop) return;
    const int j = a.step, nvec = j + 1;
    if (blockIdx.x == 0 && threadIdx.x == 0) a.flags[5] += 1;
    __shared__ int s_trig;
    __shared__ double s_max[kRoBlock / 64];
    if (a.mode == 2) {
        if (threadIdx.x == 0) s_trig = 1;
        __syncthreads();
    } else {
        const double bn = sqrt(sum_partials<kRoBlock>(a.ax_part, a.ax_n));
        double *wn = a.omega + (size_t)((j + 1) % 3) * a.ld;
        const double *wc = a.omega + (size_t)(j % 3) * a.ld;
        const double *wp = a.omega + (size_t)((j + 2) % 3) * a.ld;
        // |T| estimate: Gershgorin bound of the tridiagonal so far
        double tn = 0.0;
        for (int k = threadIdx.x; k <= j; k += kRoBlock)
            tn = fmax(tn, fabs(a.alpha[k]) + a.beta[k] + (k < j ? a.beta[k + 1] : bn));
        const double psi = a.eps * block_max<kRoBlock>(tn, s_max);
        // the step after a triggered one reorthogonalizes too; workgroup 0
        // writes that decision into the other parity's slot, so that no
        // workgroup of this launch reads a value written in it
        const bool force = a.flags[3 + (j & 1)] != 0;
        const bool lead = blockIdx.x == 0;
        double mx = 0.0;
        for (int k = threadIdx.x; k < j; k += kRoBlock) {
            double t = a.beta[k + 1] * wc[k + 1] + (a.alpha[k] - a.alpha[j]) * wc[k] - a.beta[j] * wp[k];
            if (k > 0) t += a.beta[k] * wc[k - 1];
            t += t >= 0.0 ? psi : -psi;
            const double w = bn > 0.0 ? t / bn : 0.0;
            if (lead) wn[k] = w;
            mx = fmax(mx, fabs(w));
        }
        mx = block_max<kRoBlock>(mx, s_max);
        const bool trig = force || mx > sqrt(a.eps);
        if (threadIdx.x == 0) s_trig = trig;
        if (lead) {  // the new omega row, reset to eps after a reorthogonalization
            if (trig)
                for (int k = threadIdx.x; k < j; k += kRoBlock) wn[k] = a.eps;
            if (threadIdx.x == 0) {
                wn[j] = trig ? a.eps : (bn > 0.0 ? psi / bn : 0.0);
                wn[j + 1] = 1.0;
                a.flags[0] = trig;
                a.flags[3 + ((j + 1) & 1)] = trig && !force;
            }
        }
        __syncthreads();
    }
    if (a.mode == 2 && blockIdx.x == 0 && threadIdx.x == 0) a.flags[0] = 1;
    if (!s_trig) return;
    // partial dots <u_k, r> over this workgroup's chunk of 16-byte vectors,
    // kRoGroup stored vectors per pass (their loads in flight together; the
    // chunk of r is re-read from L2 by the later passes)
    constexpr int V = 16 / sizeof(T);
    using Vec = Row<T, V>;
    const long long nv = n / V;
    const long long chunk = (nv + gridDim.x - 1) / gridDim.x;
    const long long c0 = (long long)blockIdx.x * chunk, c1 = c0 + chunk < nv ? c0 + chunk : nv;
    const Vec *rv = reinterpret_cast<const Vec *>(r);
    const int ndot = gridDim.x;
    for (int k0 = 0; k0 < nvec; k0 += kRoGroup) {
        const int nk = nvec - k0 < kRoGroup ? nvec - k0 : kRoGroup;
        double acc[kRoGroup] = {};
        for (long long q = c0 + threadIdx.x; q < c1; q += kRoBlock) {
            const Vec x = rv[q];
            Vec u[kRoGroup];
#pragma unroll
            for (int g = 0; g < kRoGroup; ++g)
                if (g < nk) u[g] = reinterpret_cast<const Vec *>(base + (long long)(k0 + g) * stride)[q];
#pragma unroll
            for (int g = 0; g < kRoGroup; ++g)
                if (g < nk)
#pragma unroll
                    for (int i = 0; i < V; ++i) acc[g] += (double)x.v[i] * (double)u[g].v[i];
        }
        if (blockIdx.x == 0)  // the n % V tail
            for (long long i = nv * V + threadIdx.x; i < n; i += kRoBlock)
#pragma unroll
                for (int g = 0; g < kRoGroup; ++g)
                    if (g < nk) acc[g] += (double)r[i] * (double)base[(long long)(k0 + g) * stride + i];
        double tot[kRoGroup];
        block_sums<kRoGroup, kRoBlock>(acc, tot);
        if (threadIdx.x == 0)
#pragma unroll
            for (int g = 0; g < kRoGroup; ++g)
                if (g < nk) a.dpart[(size_t)(k0 + g) * ndot + blockIdx.x] = tot[g];
    }
}

// c_k = s_k^2 <u_k, r>: workgroup k sums the ndot partials of vector k in a
// fixed order.
__global__ __launch_bounds__(kRoBlock) void k_ro_reduce(ReorthArgs a, int ndot)
{
    if (*a.stop || !a.flags[0]) return;
    const int k = blockIdx.x;
    double v[1] = {0.0}, h[1];
    for (int g = threadIdx.x; g < ndot; g += kRoBlock) v[0] += a.dpart[(size_t)k * ndot + g];
    block_sums<1, kRoBlock>(v, h);
    if (threadIdx.x == 0) {
        const double sk = a.scale[k];
        a.coef[k] = sk * sk * h[0];
        if (k == 0) a.flags[2] += 1;
    }
}

// r <- r - sum_k c_k u_k in FP64 on 16-byte vectors, kRoGroup stored
// vectors' loads in flight together; the |r|^2 partials of the result
// replace the axpy's (one per workgroup, gridDim.x == the axpy's count).
template <typename T>
__global__ __launch_bounds__(kRoUpdBlock) void k_ro_update(ReorthArgs a, T *r, const T *base, long long stride,
                                                           long long n)
{
    if (*a.stop || !a.flags[0]) return;
    const int nvec = a.step + 1;
    constexpr int V = 16 / sizeof(T);
    using Vec = Row<T, V>;
    const long long nv = n / V;
    Vec *rv = reinterpret_cast<Vec *>(r);
    double v[1] = {0.0};
    for (long long q = (long long)blockIdx.x * kRoUpdBlock + threadIdx.x; q < nv;
         q += (long long)gridDim.x * kRoUpdBlock) {
        const Vec x = rv[q];
        double acc[V];
#pragma unroll
        for (int i = 0; i < V; ++i) acc[i] = (double)x.v[i];
        for (int k0 = 0; k0 < nvec; k0 += kRoGroup) {
            const int nk = nvec - k0 < kRoGroup ? nvec - k0 : kRoGroup;
            Vec u[kRoGroup];
#pragma unroll
            for (int g = 0; g < kRoGroup; ++g)
                if (g < nk) u[g] = reinterpret_cast<const Vec *>(base + (long long)(k0 + g) * stride)[q];
#pragma unroll
            for (int g = 0; g < kRoGroup; ++g)
                if (g < nk) {
                    const double c = a.coef[k0 + g];
#pragma unroll
                    for (int i = 0; i < V; ++i) acc[i] -= c * (double)u[g].v[i];
                }
        }
        Vec w;
#pragma unroll
        for (int i = 0; i < V; ++i) {
            w.v[i] = (T)acc[i];
            v[0] += (double)w.v[i] * (double)w.v[i];
        }
        rv[q] = w;
    }
    if (blockIdx.x == 0)
        for (long long i = nv * V + threadIdx.x; i < n; i += kRoUpdBlock) {
            double acc = (double)r[i];
            for (int k = 0; k < nvec; ++k) acc -= a.coef[k] * (double)base[(long long)k * stride + i];
            const T t = (T)acc;
            r[i] = t;
            v[0] += (double)t * (double)t;
        }
    block_partial<1, kRoUpdBlock>(v, a.ax_part);
}

template <typename T>
int launch_reorth(hipStream_t s, const ReorthArgs &a, T *r, const T *base, long long stride, long long n,
                  int ndot)
{
    if (((uintptr_t)r | (uintptr_t)base | (uintptr_t)(stride * (long long)sizeof(T))) & 15) return -1;
    hipLaunchKernelGGL(k_ro_dot<T>, dim3(ndot), dim3(kRoBlock), 0, s, a, (const T *)r, base, stride, n);
    if (hipGetLastError() != hipSuccess) return -1;
    hipLaunchKernelGGL(k_ro_reduce, dim3(a.step + 1), dim3(kRoBlock), 0, s, a, ndot);
    if (hipGetLastError() != hipSuccess) return -1;
    hipLaunchKernelGGL(k_ro_update<T>, dim3(a.ax_n), dim3(kRoUpdBlock), 0, s, a, r, base, stride, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
template int launch_reorth<float>(hipStream_t, const ReorthArgs &, float *, const float *, long long, long long, int);
template int launch_reorth<double>(hipStream_t, const ReorthArgs &, double *, const double *, long long, long long,
                                   int);

}  // namespace edg
