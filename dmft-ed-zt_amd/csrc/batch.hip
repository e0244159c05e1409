// Batched Lanczos steps over many independent chains (uneven sector sizes)
// of single-orbital and two-orbital impurities.  Launches of k_bhv
// (single-orbital chains with short columns, several columns per tile), of
// k_bhvw (long columns, the single-chain tile) and of k_bhv2 (two-orbital
// chains, the single-chain Kanamori tile) run the H.v half of step j_i for
// every chain i, one launch of k_baxpy the axpy
// half, one launch of k_badvance moves every chain to its next step; each
// chain keeps its own step index, alpha/beta arrays and partial sums, so
// chains at different steps (admitted at different times) share the
// launches.  The recurrence is the one of lanczos_solver.f90:92-173 in the
// unnormalized-vector form used by the single-chain path:
//   H.v:   y = s_j H u_j - beta_j s_{j-1} u_{j-1},  partial <u_j, y>
//   axpy:  alpha_j = s_j <u_j, y>,  y <- y - alpha_j s_j u_j,  partial |y|^2
// with beta_{j+1} = |y| and s_{j+1} = 1/beta_{j+1} taken by the next H.v.
// A chain's reductions depend only on its own tiles (fixed count and
// order), so its result is the same whatever else is in the batch.
#include "device_util.h"

namespace edg {


template <typename T>
__device__ __forceinline__ T *slot_of(const BInst &I, int k)
{
    return static_cast<T *>(I.base) + (long long)(k % I.nslot) * I.stride;
}

// H.v tile: columns [c0, c0 + ncol) of chain i, every row.  The tile's
// columns are contiguous in memory and staged in LDS; up hops gather from
// the staged column, down hops read the neighbour column at the same row
// from global memory (coalesced over the rows).  Hop records are the 32-bit
// compressed tables (target rank | value index << 24, KC slots per row,
// self-padded).  Each tile column's neighbour-column byte offsets and
// elements are decoded once into LDS, and the vectors are read and written
// through buffer resources with 32-bit offsets: a down hop then costs one
// vector add (a 64-bit global address cost four instructions per hop).
template <typename T, typename ACC, int KS>
__global__ __launch_bounds__(kBlock) void k_bhv(const BInst *__restrict__ inst, const int *__restrict__ hv_off, const int *__restrict__ own,
                       int sub)
{
    extern __shared__ __align__(16) unsigned char smem[];
    T *xs = reinterpret_cast<T *>(smem);
    __shared__ uint32_t s_doff[kBatchMaxCols * KS];  // neighbour-column byte offsets
    __shared__ ACC s_dv[kBatchMaxCols * KS];         // their elements
    __shared__ ACC s_ed[kBatchMaxCols];
    __shared__ uint8_t s_md[kBatchMaxCols];
    __shared__ ACC s_vu[2 * (kMaxHops + 1)];

    const int i = own[blockIdx.x];  // chain of this workgroup
    const BInst &I = inst[i];
    const int t = blockIdx.x - hv_off[i];
    if (I.flags[0] || t >= I.hvn) return;  // stopped chain, or XCD padding tile
    const HvOp op = I.op;  // private copy: stores to y cannot alias it
    const int nup = op.nup, ndn = op.ndn, cb = I.cb;
    // tile t runs on XCD t % 8: contiguous column ranges per XCD keep the
    // down-hop neighbour columns in that XCD's L2
    const int c0 = xcd_group(t, I.hvn) * cb;
    const int ncol = ndn - c0 < cb ? ndn - c0 : cb;
    constexpr int kcd = KS;
    const int j = uniform(*I.step) + sub;
    const T *x = uniform(slot_of<T>(I, j));
    T *y = uniform(slot_of<T>(I, j + 1));
    const T *pv = j > 0 ? uniform(slot_of<T>(I, j - 1)) : nullptr;

    for (int e = threadIdx.x; e <= kMaxHops; e += kBlock) {
        const ACC vu = e < op.nvup ? (ACC)gload(op.up.vsrc + e) : ACC(0);
        s_vu[e] = vu;
        s_vu[kMaxHops + 1 + e] = -vu;
    }
    const int nel = ncol * nup;
    const T *xt = x + (long long)c0 * nup;
    // the tile's columns are contiguous: eight loads in flight per thread
    // before the LDS stores (one at a time measured 3% fewer solves/s on a
    // batch of 512 problems with Ns in {8, 10, 12})
    for (int e0 = threadIdx.x; e0 < nel; e0 += 8 * kBlock) {
        T v[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int e = e0 + r * kBlock;
            v[r] = gload(xt + (e < nel ? e : nel - 1));
        }
#pragma unroll
        for (int r = 0; r < 8; ++r)
            if (e0 + r * kBlock < nel) xs[e0 + r * kBlock] = v[r];
    }
    for (int e = threadIdx.x; e < ncol * kcd; e += kBlock) {
        // value index vi: [vi] = +v, [kMaxHops + 1 + vi] = -v of the signed
        // hybridization v = vsrc[vi] (vi >= nvup: a zero element)
        const uint32_t w = gload(op.dn.hrec + (long long)c0 * kcd + e);
        const uint32_t vi = rec_vidx(w), vb = vi > (uint32_t)kMaxHops ? vi - kMaxHops - 1 : vi;
        const ACC v = vb < (uint32_t)op.nvup ? (ACC)gload(op.dn.vsrc + vb) : ACC(0);
        s_doff[e] = rec_rank(w) * (uint32_t)nup * (uint32_t)sizeof(T);
        s_dv[e] = vi > (uint32_t)kMaxHops ? -v : v;
    }
    for (int c = threadIdx.x; c < ncol; c += kBlock) {
        s_ed[c] = energy_of<ACC>(op.dn, c0 + c);
        s_md[c] = gload(op.dn.mask + c0 + c);
    }
    __syncthreads();

    HvArgs la{};
    la.step = j;
    la.prev_part = I.axpart;
    la.prev_n = I.axn;
    la.scale = I.scale;
    la.beta = I.beta;
    la.stop = I.flags;
    la.nstep = I.flags + 1;
    double sx, bsp;
    if (!lanczos_scales<kBlock>(la, sx, bsp, t == 0)) return;

    // thread -> (row, column group): columns of at most kBlock rows are
    // covered by G = kBlock / nup groups of nup threads (group g takes columns
    // g, g + G, ...); longer columns by one group striding over the rows.
    // A thread loads its row's hop record, energy and mask once for all the
    // columns it handles.
    const ACC U = (ACC)op.U;
    const int span = nup <= kBlock ? nup : kBlock;
    const int G = kBlock / span;
    const int g = threadIdx.x / span;
    double dot = 0.0;
    const long long vbytes = (long long)nup * ndn * (long long)sizeof(T);
    const __amdgpu_buffer_rsrc_t xrs = buf_rsrc(x, vbytes), yrs = buf_rsrc(y, vbytes);
    const __amdgpu_buffer_rsrc_t prs = buf_rsrc(pv ? pv : x, vbytes);
    const double bq = pv ? bsp : 0.0;
    if (g < G) {
        for (int iup = threadIdx.x - g * span; iup < nup; iup += span) {
            const Row<uint32_t, KS> ur = gload(reinterpret_cast<const Row<uint32_t, KS> *>(op.up.hrec + (long long)iup * KS));
            const ACC eu = energy_of<ACC>(op.up, iup);
            const bool mu = gload(op.up.mask + iup) != 0;
            const uint32_t ib = (uint32_t)iup * (uint32_t)sizeof(T);
            for (int c = g; c < ncol; c += G) {
                const T *xc = xs + c * nup;
                const uint32_t *doff = s_doff + c * KS;
                const ACC *dv = s_dv + c * KS;
                T gn[KS];
#pragma unroll
                for (int k = 0; k < KS; ++k) gn[k] = bload<T>(xrs, doff[k] + ib, 0);
                const uint32_t eb = (uint32_t)(c0 + c) * (uint32_t)nup * (uint32_t)sizeof(T) + ib;
                // step 0 (no u_{j-1}) reads x here and weights it by 0: no
                // branch, so nothing waits on the loads before their use
                const T old = bload<T>(prs, eb, 0);
                const ACC xv = (ACC)xc[iup];
                ACC acc = (eu + s_ed[c] + (mu && s_md[c] ? U : ACC(0))) * xv;
#pragma unroll
                for (int k = 0; k < KS; ++k) acc += s_vu[rec_vidx(ur.v[k])] * (ACC)xc[rec_rank(ur.v[k])];
#pragma unroll
                for (int k = 0; k < KS; ++k) acc += dv[k] * (ACC)gn[k];
                const T rt = (T)(sx * (double)acc - bq * (double)old);
                bstore<T>(yrs, eb, rt);
                dot += (double)xv * (double)rt;
            }
        }
    }
    double mine[1] = {dot};
    block_partial_at<1>(mine, I.hvpart, t);
}

// H.v tile for chains with long sector columns: the single-chain tile
// (hv1_body.inc: staged columns for the up hops, wave-uniform neighbour-column
// loads for the down hops) over C columns per 512-thread workgroup.  Every
// chain's tiles start at a multiple of 8 in the launch, so tile t runs on
// XCD t % 8 and xcd_group deals the chain's column groups to the XCDs in
// contiguous ranges, as in the single-chain kernel; tiles past the chain's
// hvn are padding and exit at once.
template <typename T, int C, int SLOTS>
__global__ __launch_bounds__(kHvBlock) void k_bhvw(const BInst *__restrict__ inst, const int *__restrict__ hv_off, const int *__restrict__ own,
                       int sub)
{
    const int i = own[blockIdx.x];  // chain of this workgroup
    const BInst &I = inst[i];
    const int t = blockIdx.x - hv_off[i];
    if (I.flags[0] || t >= I.hvn) return;
    const int j = uniform(*I.step) + sub;
    HvArgs la{};
    la.x = uniform(slot_of<T>(I, j));
    la.y = uniform(slot_of<T>(I, j + 1));
    la.p = j > 0 ? uniform(slot_of<T>(I, j - 1)) : nullptr;
    la.partial = I.hvpart;
    la.step = j;
    la.prev_part = I.axpart;
    la.prev_n = I.axn;
    la.scale = I.scale;
    la.beta = I.beta;
    la.stop = I.flags;
    la.nstep = I.flags + 1;
    // a private copy: the tile's stores to y cannot alias it, so its fields
    // stay in registers across the row loop
    const HvOp op = I.op;
    const HvArgs &a = la;
    using ACC = T;
    constexpr int MODE = HV_LANCZOS, KH = SLOTS;
    const int idn0 = xcd_group(t, I.hvn) * C;
    const int slot = t;
    const bool lead = t == 0;
#include "hv1_body.inc"
}

// H.v tile of two-orbital (Kanamori) chains: the single-problem two-orbital
// tile (hv2_body.inc: the C columns and their orbital-move target columns
// staged in LDS, neighbour-column loads for the down hops) over C columns
// per 512-thread workgroup, XCD-aligned tile offsets as in k_bhvw.
template <typename T, typename ACC, int C, int SLOTS>
__global__ __launch_bounds__(kHvBlock) void k_bhv2(const BInst *__restrict__ inst, const int *__restrict__ hv_off,
                                                   const int *__restrict__ own, int sub)
{
    const int i = own[blockIdx.x];  // chain of this workgroup
    const BInst &I = inst[i];
    const int t = blockIdx.x - hv_off[i];
    if (I.flags[0] || t >= I.hvn) return;
    const int j = uniform(*I.step) + sub;
    HvArgs la{};
    la.x = uniform(slot_of<T>(I, j));
    la.y = uniform(slot_of<T>(I, j + 1));
    la.p = j > 0 ? uniform(slot_of<T>(I, j - 1)) : nullptr;
    la.partial = I.hvpart;
    la.step = j;
    la.prev_part = I.axpart;
    la.prev_n = I.axn;
    la.scale = I.scale;
    la.beta = I.beta;
    la.stop = I.flags;
    la.nstep = I.flags + 1;
    const HvOp op = I.op;
    const HvArgs &a = la;
    constexpr int MODE = HV_LANCZOS, KH = SLOTS;
    const int idn0 = xcd_group(t, I.hvn) * C;
    const int slot = t;
    const bool lead = t == 0;
#include "hv2_body.inc"
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_baxpy(const BInst *__restrict__ inst, const int *__restrict__ ax_off, const int *__restrict__ own,
                        int sub)
{
    const int i = own[blockIdx.x];  // chain of this workgroup
    const BInst &I = inst[i];
    if (I.flags[0]) return;
    const int t = blockIdx.x - ax_off[i];
    const int j = uniform(*I.step) + sub;
    const double sj = I.scale[j];
    const double al = sj * sum_partials<kBlock>(I.hvpart, I.hvn);
    if (t == 0 && threadIdx.x == 0) I.alpha[j] = al;
    const double coef = al * sj;
    T *r = uniform(slot_of<T>(I, j + 1));
    const T *u = uniform(slot_of<T>(I, j));
    // 16-byte vectors (slots are 256-byte aligned); workgroup 0 also takes
    // the n % V tail
    constexpr int V = 16 / sizeof(T);
    using Vec = Row<T, V>;
    const long long n = (long long)I.op.nup * I.op.ndn, nv = n / V;
    const long long stride = (long long)I.axn * kBlock;
    double v[1] = {0.0};
    for (long long q = (long long)t * kBlock + threadIdx.x; q < nv; q += stride) {
        const Vec rv = gload(reinterpret_cast<const Vec *>(r) + q);
        const Vec uv = gload(reinterpret_cast<const Vec *>(u) + q);
        Vec w;
#pragma unroll
        for (int i = 0; i < V; ++i) {
            w.v[i] = (T)((double)rv.v[i] - coef * (double)uv.v[i]);
            v[0] += (double)w.v[i] * (double)w.v[i];
        }
        gstore(reinterpret_cast<Vec *>(r) + q, w);
    }
    if (t == 0)
        for (long long e = nv * V + threadIdx.x; e < n; e += kBlock) {
            const T w = (T)((double)gload(r + e) - coef * (double)gload(u + e));
            gstore(r + e, w);
            v[0] += (double)w * (double)w;
        }
    block_partial_at<1>(v, I.axpart, t);
}

__global__ void k_badvance(const BInst *inst, int ninst, int k)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ninst) *inst[i].step += k;
}

template <int KS>
static int set_lds_ks(size_t max_lds)
{
    const void *fs[6] = {(const void *)k_bhv<float, float, KS>,     (const void *)k_bhv<double, double, KS>,
                         (const void *)k_bhvw<float, 1, KS>,        (const void *)k_bhvw<double, 1, KS>,
                         (const void *)k_bhvw<float, kBatchWideCols, KS>,
                         (const void *)k_bhvw<double, kBatchWideCols, KS>};
    for (const void *f : fs)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_lds) != hipSuccess)
            return -1;
    return 0;
}

template <int KS>
static int set_lds_kan(size_t max_lds)
{
    const void *fs[2] = {(const void *)k_bhv2<float, float, kBatchKanColsF32, KS>,
                         (const void *)k_bhv2<double, double, 1, KS>};
    for (const void *f : fs)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_lds) != hipSuccess)
            return -1;
    return 0;
}

int batch_set_lds(size_t max_lds)
{
    return set_lds_ks<8>(max_lds) || set_lds_ks<16>(max_lds) || set_lds_kan<8>(max_lds) ||
                   set_lds_kan<12>(max_lds) || set_lds_kan<16>(max_lds)
               ? -1
               : 0;
}

template <int C>
static void (*wide_kernel(bool fp64, int ks))(const BInst *, const int *, const int *, int)
{
    if (ks == 8) return fp64 ? k_bhvw<double, C, 8> : k_bhvw<float, C, 8>;
    return fp64 ? k_bhvw<double, C, 16> : k_bhvw<float, C, 16>;
}

int launch_batch_hv(hipStream_t s, bool fp64, int wide_cols, int ks, const BInst *inst, const int *hv_off,
                    const int *own, int ntiles, int lds_elems, int sub)
{
    if (ntiles <= 0) return 0;
    const size_t lds = (size_t)lds_elems * (fp64 ? 8 : 4);
    const bool wide = wide_cols > 0;
    void (*f)(const BInst *, const int *, const int *, int);
    if (wide) {
        f = wide_cols == 1 ? wide_kernel<1>(fp64, ks) : wide_kernel<kBatchWideCols>(fp64, ks);
    } else {
        if (ks == 8) f = fp64 ? k_bhv<double, double, 8> : k_bhv<float, float, 8>;
        else f = fp64 ? k_bhv<double, double, 16> : k_bhv<float, float, 16>;
    }
    hipLaunchKernelGGL(f, dim3(ntiles), dim3(wide ? kHvBlock : kBlock), lds, s, inst, hv_off, own, sub);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_batch_hv2(hipStream_t s, bool fp64, int kh, const BInst *inst, const int *hv_off, const int *own,
                     int ntiles, size_t lds_bytes, int sub)
{
    if (ntiles <= 0) return 0;
    void (*f)(const BInst *, const int *, const int *, int);
    if (kh == 8) f = fp64 ? k_bhv2<double, double, 1, 8> : k_bhv2<float, float, kBatchKanColsF32, 8>;
    else if (kh == 12) f = fp64 ? k_bhv2<double, double, 1, 12> : k_bhv2<float, float, kBatchKanColsF32, 12>;
    else if (kh == 16) f = fp64 ? k_bhv2<double, double, 1, 16> : k_bhv2<float, float, kBatchKanColsF32, 16>;
    else return -1;
    hipLaunchKernelGGL(f, dim3(ntiles), dim3(kHvBlock), lds_bytes, s, inst, hv_off, own, sub);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_batch_axpy(hipStream_t s, bool fp64, const BInst *inst, const int *ax_off, const int *own,
                      int nblocks, int sub)
{
    if (nblocks <= 0) return 0;
    if (fp64) hipLaunchKernelGGL(k_baxpy<double>, dim3(nblocks), dim3(kBlock), 0, s, inst, ax_off, own, sub);
    else hipLaunchKernelGGL(k_baxpy<float>, dim3(nblocks), dim3(kBlock), 0, s, inst, ax_off, own, sub);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// A chain's state block zeroed, with its flags (the two ints at byte 8)
// set to {0, maxn}.
__global__ void k_state_init(unsigned long long *st, long long n8, int maxn)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n8; i += (long long)gridDim.x * blockDim.x)
        st[i] = i == 1 ? (unsigned long long)(unsigned)maxn << 32 : 0ull;
}

int launch_batch_state_init(hipStream_t s, void *st, size_t bytes, int maxn)
{
    const long long n8 = (long long)(bytes / 8);
    const int grid = (int)std::min<long long>((n8 + 255) / 256, 64);
    hipLaunchKernelGGL(k_state_init, dim3(grid), dim3(256), 0, s, (unsigned long long *)st, n8, maxn);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_batch_advance(hipStream_t s, const BInst *inst, int ninst, int k)
{
    if (ninst <= 0) return 0;
    hipLaunchKernelGGL(k_badvance, dim3((ninst + 255) / 256), dim3(256), 0, s, inst, ninst, k);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace edg
