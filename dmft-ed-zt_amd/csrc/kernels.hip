// HIP kernels of the ED impurity solver for gfx950 (MI355X).
//
// Sector vectors are stored as an nup x ndn column-major matrix X (up rank
// fastest), the row order of ed_basis.f90:132-135.  The Hamiltonian of
// impurity_hamiltonian.f90:77-247 splits as
//     H = D(up,dn) + H_up (x) 1 + 1 (x) H_dn + K
// where H_up / H_dn are the per-spin impurity-bath hoppings (same matrix for
// every column / row), D is diagonal (one-spin energies plus the U / U'
// cross terms) and K the Kanamori spin-flip and pair-hopping terms (norb>=2).
// One workgroup owns C adjacent columns: it stages them in LDS with the C
// values of a row interleaved, so an up-hop is one wide LDS read serving all
// C columns, while down-hops read whole neighbour columns from global memory
// with coalesced wave64 loads.  Reductions are FP64 and deterministic: each
// block writes a partial, and the consuming kernel sums them in a fixed order.
#include "device_util.h"

namespace edg {

enum FinMode { FIN_ALPHA = 0, FIN_BETA = 1, FIN_NORM = 2, FIN_COPY = 3, FIN_OBS = 4, FIN_FUSED = 5 };

struct FinArgs {
    int mode;
    int nblk;
    const double *sx;   // FIN_ALPHA: alpha = sx * sum
    double *out0;       // alpha / beta / norm / sums (FIN_FUSED: the alpha array)
    double *out1;       // scale (1/beta, 1/norm) (FIN_FUSED: the beta array)
    int *stop;          // FIN_BETA / FIN_FUSED: set when beta < 1e-16; FIN_ALPHA/BETA/FUSED skip when set
    int *nstep;
    int step;
    int norb;           // FIN_OBS
    double *out2;       // FIN_FUSED: the scale array
    FusedCoef *fc;      // FIN_FUSED: the coefficients of the next step go to fc[step + 1]
};

__device__ __forceinline__ void put_coef(FusedCoef *fc, const double (&v)[7])
{
#pragma unroll
    for (int i = 0; i < 7; ++i) {
        fc->d[i] = v[i];
        fc->f[i] = (float)v[i];
    }
    fc->d[7] = 0.0;
    fc->f[7] = 0.0f;
}

constexpr int kFinBlock = 1024;

template <int NS>
__global__ __launch_bounds__(kFinBlock) void k_finish(const double *partial, FinArgs f)
{
    if ((f.mode == FIN_ALPHA || f.mode == FIN_BETA || f.mode == FIN_FUSED) && f.stop && *f.stop) return;
    double v[NS], tot[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) v[s] = 0.0;
    for (int b = threadIdx.x; b < f.nblk; b += kFinBlock)
#pragma unroll
        for (int s = 0; s < NS; ++s) v[s] += partial[(long long)b * NS + s];
    block_sums<NS, kFinBlock>(v, tot);
    if (threadIdx.x != 0) return;
    switch (f.mode) {
    case FIN_ALPHA:
        f.out0[0] = (*f.sx) * tot[0];
        break;
    case FIN_BETA: {
        const double b = sqrt(tot[0]);
        f.out0[0] = b;
        f.out1[0] = b > 0.0 ? 1.0 / b : 0.0;
        if (b < 1.0e-16) {
            *f.stop = 1;
            *f.nstep = f.step + 1;
        }
        break;
    }
    case FIN_NORM: {
        const double nr = sqrt(tot[0]);
        if (f.out0) f.out0[0] = nr;
        if (f.out1) f.out1[0] = 1.0 / nr;
        break;
    }
    case FIN_COPY:
#pragma unroll
        for (int s = 0; s < NS; ++s) f.out0[s] = tot[s];
        break;
    case FIN_FUSED:
        if constexpr (NS >= 3) {
            // tot = <u_j, r_j>, |r_j|^2, |u_j|^2 for the residual shifted by
            // sigma_j = alpha_{j-1} (0 at j = 0), r_j = (H - sigma_j) q_j -
            // beta_j q_{j-1}, q_j = s_j u_j: the reference's alpha_j =
            // <q_j, H q_j - beta_j q_{j-1}> = s_j tot0 + sigma_j |q_j|^2 and
            // beta_{j+1}^2 = |r_j - d q_j|^2 with d = alpha_j - sigma_j
            const double sj = f.out2[f.step];
            const double sg = f.step > 0 ? f.out0[f.step - 1] : 0.0;
            const double qq = sj * sj * tot[2];
            const double al = sj * tot[0] + sg * qq;
            const double d = al - sg;
            const double b2 = tot[1] - 2.0 * d * sj * tot[0] + d * d * qq;
            const double b = b2 > 0.0 ? sqrt(b2) : 0.0;
            f.out0[f.step] = al;
            f.out1[f.step + 1] = b;
            const double sn = b > 0.0 ? 1.0 / b : 0.0;
            f.out2[f.step + 1] = sn;
            if (b < 1.0e-16) {
                *f.stop = 1;
                *f.nstep = f.step + 1;
            }
            // step j+1: s_{j+1}, d = alpha_j - sigma_j, d s_j, sigma_j s_j,
            // beta_j s_{j-1}, sigma_{j+1} s_{j+1} (sigma_{j+1} = alpha_j), beta_{j+1} s_j
            const double bj = f.out1[f.step], sjm1 = f.step > 0 ? f.out2[f.step - 1] : 0.0;
            const double v[7] = {sn, d, d * sj, sg * sj, bj * sjm1, al * sn, b * sj};
            put_coef(f.fc + f.step + 1, v);
        }
        break;
    default:  // FIN_OBS: sums laid out [m], [kMaxNorb+m], [2 kMaxNorb+m]
        for (int m = 0; m < f.norb; ++m) {
            f.out0[m] = tot[m];
            f.out0[f.norb + m] = tot[kMaxNorb + m];
            f.out0[2 * f.norb + m] = tot[2 * kMaxNorb + m];
        }
    }
}

template <int NS>
static int launch_finish(hipStream_t s, const double *partial, const FinArgs &f)
{
    hipLaunchKernelGGL(k_finish<NS>, dim3(1), dim3(kFinBlock), 0, s, partial, f);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Finisher of the H.v reductions: alpha for HV_LANCZOS, the three sums for HV_RQ.
static int finish_hv(hipStream_t s, int mode, int nblk, const HvArgs &a)
{
    if (mode == HV_RQ)
        if (launch_finish<3>(s, a.partial, FinArgs{FIN_COPY, nblk, nullptr, a.result, nullptr,
                                                   nullptr, nullptr, 0, 0, nullptr, nullptr}))
            return -1;
    return nblk;
}

int launch_alpha(hipStream_t s, const double *partial, int nblk, const double *scale_j,
                 double *alpha_j, int *stop)
{
    return launch_finish<1>(s, partial, FinArgs{FIN_ALPHA, nblk, scale_j, alpha_j, nullptr, stop,
                                                nullptr, 0, 0, nullptr, nullptr});
}

int launch_fused_finish(hipStream_t s, const double *partial, int nblk, int step, double *alpha,
                        double *beta, double *scale, int *stop, int *nstep, FusedCoef *fc)
{
    return launch_finish<3>(s, partial, FinArgs{FIN_FUSED, nblk, nullptr, alpha, beta, stop, nstep, step,
                                                0, scale, fc});
}

__global__ void k_fused_start(const double *scale, FusedCoef *fc)
{
    if (threadIdx.x == 0) {
        const double v[7] = {scale[0], 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        put_coef(fc, v);
    }
}

int launch_fused_start(hipStream_t s, const double *scale, FusedCoef *fc)
{
    hipLaunchKernelGGL(k_fused_start, dim3(1), dim3(64), 0, s, scale, fc);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

__device__ __forceinline__ uint32_t binom_tab(const uint32_t (*B)[16], int n, int k)
{
    return (k < 0 || k > n) ? 0u : B[n][k];
}

// ------------------------------------------------------ per-spin tables

// Compressed [n][KC] copies of row r's first KC slots: 16-bit target rank
// and 8-bit element index (bit 7 = odd parity, the element being minus the
// signed V at that index; 64 = zero element); and the same pair as one
// 32-bit record, rank | (value-table index << 24).  Reads the row's full
// records, so it runs after them in the same thread or a later launch.
__device__ inline void compress_row(const SpinTab &t, int r)
{
    for (int k = 0; k < t.KC; ++k) {
        const HopD h = t.hd[(long long)r * t.K + k];
        t.ccol[(long long)r * t.KC + k] = (uint16_t)h.col;
        t.cidx[(long long)r * t.KC + k] = (uint8_t)h.pad;
        const uint32_t vi = (h.pad & 0x7Fu) + ((h.pad >> 7) & 1u ? (uint32_t)kMaxHops + 1u : 0u);
        t.hrec[(long long)r * t.KC + k] = (h.col & 0xFFFFu) | (vi << 24);
    }
}

// One thread per configuration rank r of N electrons on nsite sites.
// Rank order = increasing bit-pattern order (ed_basis.f90:94-112), i.e. the
// combinatorial number system r = sum_k C(p_k, k) over set bits p_1<..<p_N.
__global__ __launch_bounds__(kBlock) void k_spin_table(int nsite, int norb, int nbath, int ne,
                                                       const double *em_s, const double *ek_s,
                                                       const double *v_s, double mu, double UpmJ,
                                                       SpinTab t, int compress, uint32_t *cmap,
                                                       int cmap_pm)
{
    __shared__ uint32_t B[16][16];
    for (int e = threadIdx.x; e < 256; e += kBlock) {
        int n = e >> 4, k = e & 15;
        uint32_t c = 0;
        if (k <= n) {
            c = 1;
            for (int i = 1; i <= k; ++i) c = c * (uint32_t)(n - k + i) / (uint32_t)i;
        }
        B[n][k] = c;
    }
    __syncthreads();
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= t.n) return;
    // unrank
    uint32_t p = 0, rem = (uint32_t)r;
    for (int k = ne; k >= 1; --k) {
        int c = k - 1;
        while (c + 1 < nsite && B[c + 1][k] <= rem) ++c;
        p |= 1u << c;
        rem -= B[c][k];
    }
    t.pattern[r] = p;
    auto rank_of = [&](uint32_t q) {
        uint32_t s = 0;
        int k = 0;
        for (int b = 0; b < nsite; ++b)
            if ((q >> b) & 1u) { ++k; s += binom_tab(B, b, k); }
        return s;
    };
    // c+ (cmap_pm 1) / c (2) source map of the N - 1 / N + 1 sector for
    // every impurity orbital (the maps of k_cmap, cmap[m * n + r])
    if (cmap)
        for (int m = 0; m < norb; ++m) {
            const uint32_t set = (p >> m) & 1u;
            uint32_t code = kInvalid;
            if ((cmap_pm == 1 && set) || (cmap_pm == 2 && !set))
                code = rank_of(p ^ (1u << m)) | ((uint32_t)(__popc(p & ((1u << m) - 1u)) & 1) << 31);
            cmap[(long long)m * t.n + r] = code;
        }
    // hybridization hops in reference loop order (iorb outer, bath inner)
    int slot = 0;
    for (int m = 0; m < norb; ++m) {
        for (int l = 0; l < nbath; ++l) {
            const double v = v_s[l * norb + m];
            if (fabs(v) < 1.0e-16) continue;
            const int i = m, j = norb + l;
            const uint32_t bi = (p >> i) & 1u, bj = (p >> j) & 1u;
            if (bi == bj) continue;
            const uint32_t q = p ^ ((1u << i) | (1u << j));
            const uint32_t between = p & ((1u << j) - 1u) & ~((2u << i) - 1u);
            const uint32_t odd = __popc(between) & 1u;
            const double val = odd ? -v : v;
            const uint32_t c = rank_of(q);
            // pad: element index, bit 7 = odd Jordan-Wigner parity (the
            // compressed tables negate the signed V_lm stored at that index)
            t.hd[(long long)r * t.K + slot] = HopD{c, (uint32_t)(l * norb + m) | (odd << 7), val};
            t.hf[(long long)r * t.K + slot] = HopF{c, (float)val};
            ++slot;
        }
    }
    t.cnt[r] = (uint8_t)slot;
    for (int k = slot; k < t.K; ++k) {
        t.hd[(long long)r * t.K + k] = HopD{(uint32_t)r, 64u, 0.0};
        t.hf[(long long)r * t.K + k] = HopF{(uint32_t)r, 0.0f};
    }
    if (t.maxcnt) atomicMax(t.maxcnt, slot);
    if (compress) compress_row(t, r);
    // one-spin diagonal energy
    double e = 0.0;
    for (int m = 0; m < norb; ++m)
        if ((p >> m) & 1u) e += em_s[m] - mu;
    for (int m = 0; m < norb; ++m)
        for (int m2 = m + 1; m2 < norb; ++m2)
            if (((p >> m) & 1u) && ((p >> m2) & 1u)) e += UpmJ;
    for (int l = 0; l < nbath; ++l)
        if ((p >> (norb + l)) & 1u) e += ek_s[l];
    t.energy[r] = e;
    t.energyf[r] = (float)e;
    t.mask[r] = (uint8_t)(p & ((1u << norb) - 1u));
    // orbital moves dst <- src (Kanamori terms), sign of modes strictly between
    if (norb >= 2) {
        for (int dst = 0; dst < norb; ++dst)
            for (int src = 0; src < norb; ++src) {
                uint32_t code = kInvalid;
                if (dst != src && ((p >> src) & 1u) && !((p >> dst) & 1u)) {
                    const int lo = dst < src ? dst : src, hi = dst < src ? src : dst;
                    const uint32_t q = p ^ ((1u << src) | (1u << dst));
                    const uint32_t between = p & ((1u << hi) - 1u) & ~((2u << lo) - 1u);
                    code = rank_of(q) | ((uint32_t)(__popc(between) & 1) << 31);
                }
                t.orb[(long long)(dst * norb + src) * t.n + r] = code;
            }
    }
}

// Compressed tables of the sectors whose hop targets fit 16 bits (k_compress,
// or k_spin_table with compress set).
__global__ __launch_bounds__(kBlock) void k_compress(SpinTab t)
{
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= t.n) return;
    compress_row(t, r);
}

int launch_compress(hipStream_t s, const SpinTab &t)
{
    hipLaunchKernelGGL(k_compress, dim3((t.n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, t);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_spin_table(hipStream_t s, int nsite, int norb, int nbath, int ne, const double *em_s,
                      const double *ek_s, const double *v_s, double mu, double UpmJ,
                      const SpinTab &t, bool compress, uint32_t *cmap, int cmap_pm)
{
    const int grid = (t.n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_spin_table, dim3(grid), dim3(kBlock), 0, s, nsite, norb, nbath, ne, em_s,
                       ek_s, v_s, mu, UpmJ, t, compress ? 1 : 0, cmap, cmap_pm);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ------------------------------------------------------------------ H.v

__device__ __forceinline__ double cross_energy(int mu, int md, double U, double Up)
{
    const int both = __popc(mu & md);
    return U * both + Up * (double)(__popc(mu) * __popc(md) - both);
}

// y = H x over the C columns [idn0, idn0+C) of the sector matrix.
//   T    storage type of the vectors, ACC arithmetic type of the row sums
//   KH   compile-time bound on the hop count of both spins (0 = any, loops)
//   MODE HV_PLAIN: y = Hx; HV_ACCUM: y += Hx; HV_LANCZOS: y = s_j Hx - beta_j s_{j-1} p
//        with alpha_j = s_j <x, y>; HV_RQ: no store, sums <x,Hx>, <Hx,Hx>, <x,x>
template <typename T, typename ACC, int C, int MODE, bool KAN, int KH>
__global__ __launch_bounds__(kHvBlock) void k_hv(HvOp op, HvArgs a)
{
    if (a.stop && *a.stop) return;
    using Hop = typename HopOf<ACC>::type;
    extern __shared__ __align__(64) unsigned char smem[];
    T *xs = reinterpret_cast<T *>(smem);
    __shared__ int s_cnt[C];
    __shared__ int s_col[C][kMaxHops];
    __shared__ ACC s_val[C][kMaxHops];
    __shared__ double s_ed[C];
    __shared__ int s_md[C];
    __shared__ uint32_t s_orb[KAN ? C : 1][KAN ? kMaxNorb * kMaxNorb : 1];

    const int nup = op.nup, ndn = op.ndn;
    const long long ld = nup;
    const int idn0 = xcd_group(blockIdx.x, (ndn + C - 1) / C) * C;
    const T *x = static_cast<const T *>(a.x);
    const Hop *dtab = HopOf<ACC>::tab(op.dn);
    const Hop *utab = HopOf<ACC>::tab(op.up);

    for (int e = threadIdx.x; e < C * kMaxHops; e += kHvBlock) {
        const int c = e / kMaxHops, k = e - c * kMaxHops;
        const int idn = idn0 + c;
        const int cnt = idn < ndn ? op.dn.cnt[idn] : 0;
        if (k == 0) {
            s_cnt[c] = cnt;
            s_ed[c] = idn < ndn ? op.dn.energy[idn] : 0.0;
            s_md[c] = idn < ndn ? op.dn.mask[idn] : 0;
        }
        if (k < cnt) {
            const Hop h = dtab[(long long)idn * op.dn.K + k];
            s_col[c][k] = (int)h.col;
            s_val[c][k] = (ACC)h.val;
        }
    }
    if constexpr (KAN) {
        const int nn = op.norb * op.norb;
        for (int e = threadIdx.x; e < C * nn; e += kHvBlock) {
            const int c = e / nn, q = e - c * nn;
            const int idn = idn0 + c;
            s_orb[c][q] = idn < ndn ? op.dn.orb[(long long)q * ndn + idn] : kInvalid;
        }
    }
    // stage the C columns, row-interleaved: xs[i*C + c] = x(i, idn0 + c)
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int idn = idn0 + c;
        if (idn < ndn) {
            const T *xc = x + (long long)idn * ld;
            for (int i = threadIdx.x; i < nup; i += kHvBlock) xs[i * C + c] = xc[i];
        } else {
            for (int i = threadIdx.x; i < nup; i += kHvBlock) xs[i * C + c] = T(0);
        }
    }
    __syncthreads();

    double sums[3] = {0.0, 0.0, 0.0};
    double sx = 1.0, bsp = 0.0;
    if constexpr (MODE == HV_LANCZOS) {
        if (!lanczos_scales<kHvBlock>(a, sx, bsp, blockIdx.x == 0)) return;
    }
    T *y = static_cast<T *>(a.y);
    const T *pv = static_cast<const T *>(a.p);

    for (int iup = threadIdx.x; iup < nup; iup += kHvBlock) {
        ACC acc[C];
        const Row<T, C> xr = *reinterpret_cast<const Row<T, C> *>(xs + iup * C);
        const double eu = op.up.energy[iup];
        const int mu = op.up.mask[iup];
#pragma unroll
        for (int c = 0; c < C; ++c)
            acc[c] = (ACC)(eu + s_ed[c] + cross_energy(mu, s_md[c], op.U, op.Up)) * (ACC)xr.v[c];
        if constexpr (KH > 0) {
            // up hops: every table slot < KH exists (tables hold >= kMinSlots
            // slots, unused ones point at the row itself with a zero element),
            // so all records, then all LDS rows, are loaded before any FMA
            Hop h[KH];
#pragma unroll
            for (int k = 0; k < KH; ++k) h[k] = utab[(long long)iup * op.up.K + k];
            Row<T, C> src[KH];
#pragma unroll
            for (int k = 0; k < KH; ++k) src[k] = *reinterpret_cast<const Row<T, C> *>(xs + h[k].col * C);
#pragma unroll
            for (int k = 0; k < KH; ++k)
#pragma unroll
                for (int c = 0; c < C; ++c) acc[c] += (ACC)h[k].val * (ACC)src[k].v[c];
            // down hops: neighbour columns from global memory; the column base
            // is wave-uniform, so each load is scalar base + lane offset
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int cnt = __builtin_amdgcn_readfirstlane(s_cnt[c]);
                T g[KH];
#pragma unroll
                for (int k = 0; k < KH; ++k) {
                    if (k < cnt) {
                        const int col = __builtin_amdgcn_readfirstlane(s_col[c][k]);
                        g[k] = x[(long long)col * ld + iup];
                    } else {
                        g[k] = T(0);
                    }
                }
#pragma unroll
                for (int k = 0; k < KH; ++k)
                    if (k < cnt) acc[c] += s_val[c][k] * (ACC)g[k];
            }
        } else {
            const int cu = op.up.cnt[iup];
            for (int k = 0; k < cu; ++k) {
                const Hop hk = utab[(long long)iup * op.up.K + k];
                const Row<T, C> src = *reinterpret_cast<const Row<T, C> *>(xs + hk.col * C);
#pragma unroll
                for (int c = 0; c < C; ++c) acc[c] += (ACC)hk.val * (ACC)src.v[c];
            }
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int cnt = s_cnt[c];
                for (int k = 0; k < cnt; ++k)
                    acc[c] += s_val[c][k] * (ACC)x[(long long)s_col[c][k] * ld + iup];
            }
        }
        if constexpr (KAN) {
            const int norb = op.norb;
            for (int i = 0; i < norb; ++i)
                for (int j = 0; j < norb; ++j) {
                    if (i == j) continue;
                    const uint32_t tu = op.up.orb[(long long)(i * norb + j) * nup + iup];
                    if (tu == kInvalid) continue;
                    const double su = (tu >> 31) ? -op.Jp : op.Jp;
                    const long long ru = tu & 0x7FFFFFFFu;
#pragma unroll
                    for (int c = 0; c < C; ++c) {
                        const uint32_t tsf = s_orb[c][j * norb + i];  // dn i -> j: spin flip
                        const uint32_t tph = s_orb[c][i * norb + j];  // dn j -> i: pair hop
                        if (tsf != kInvalid)
                            acc[c] += (ACC)((tsf >> 31) ? -su : su) *
                                      (ACC)x[(long long)(tsf & 0x7FFFFFFFu) * ld + ru];
                        if (tph != kInvalid)
                            acc[c] += (ACC)((tph >> 31) ? -su : su) *
                                      (ACC)x[(long long)(tph & 0x7FFFFFFFu) * ld + ru];
                    }
                }
        }
#pragma unroll
        for (int c = 0; c < C; ++c) {
            if (idn0 + c >= ndn) break;
            const long long e = (long long)(idn0 + c) * ld + iup;
            if constexpr (MODE == HV_PLAIN) {
                y[e] = (T)acc[c];
            } else if constexpr (MODE == HV_ACCUM) {
                y[e] = (T)((ACC)y[e] + acc[c]);
            } else if constexpr (MODE == HV_LANCZOS) {
                double r = sx * (double)acc[c];
                if (pv) r -= bsp * (double)pv[e];
                const T rt = (T)r;
                y[e] = rt;
                sums[0] += (double)xr.v[c] * (double)rt;
            } else {
                const double xv = (double)xr.v[c], hv = (double)acc[c];
                sums[0] += xv * hv;
                sums[1] += hv * hv;
                sums[2] += xv * xv;
            }
        }
    }
    if constexpr (MODE == HV_LANCZOS) {
        double mine[1] = {sums[0]};
        block_partial<1, kHvBlock>(mine, a.partial);  // alpha_j = s_j * sum (finisher)
    } else if constexpr (MODE == HV_RQ) {
        block_partial<3, kHvBlock>(sums, a.partial);
    }
}

// Single-chain single-orbital H.v: workgroup b owns the column group
// xcd_group(b) and writes its partial sums at slot b.
template <typename T, typename ACC, int C, int MODE, int KH>
__global__ __launch_bounds__(kHvBlock) void k_hv1(HvOp op, HvArgs a)
{
    if (a.stop && *a.stop) return;
    const int idn0 = xcd_group(blockIdx.x, (op.ndn + C - 1) / C) * C;
    const int slot = blockIdx.x;
    const bool lead = blockIdx.x == 0;
#include "hv1_body.inc"
}

// Two-orbital (Kanamori) H.v over the sector columns [idn0, idn0 + C), one
// 512-thread workgroup per column group (XCD-aware group order).  For each
// ordered orbital pair (i, j) the spin-flip term moves the down electron
// i -> j and the pair-hop term j -> i; with two orbitals a down pattern
// whose impurity part is (1,0) or (0,1) allows exactly one move per pair and
// both pairs lead to the same target column, (1,1) and (0,0) allow none.  So
// each column has at most one orbital-move target column, staged in LDS next
// to the column itself (zeros when there is none), and the spin-flip and
// pair-hop elements +-Jp * s_up * s_dn (impurity_hamiltonian.f90:176-230)
// gather from it at the up-move target row.  Columns are row-interleaved in
// LDS, xs[i*C + c], so a gather reads all C columns at once, and the per-row
// loads (up-hop records, up moves, mask, energy) serve all C columns.  Up
// hops use the 32-bit records (KS slots stored, KH processed); down hops are
// coalesced loads of neighbour columns at wave-uniform offsets.  The
// diagonal adds U n_m,up n_m,dn and U' for unlike orbitals of opposite spin
// to the one-spin energies (which carry U' - J for like spins).
template <typename T, typename ACC, int C, int MODE, int KH>
__global__ __launch_bounds__(kHvBlock) void k_hv2(HvOp op, HvArgs a)
{
    if (a.stop && *a.stop) return;
    const int idn0 = xcd_group(blockIdx.x, (op.ndn + C - 1) / C) * C;
    const int slot = blockIdx.x;
    const bool lead = blockIdx.x == 0;
#include "hv2_body.inc"
}

template <typename T, typename ACC, int C, int KH>
static int launch_hv2_c(hipStream_t s, const HvOp &op, int mode, const HvArgs &a)
{
    const size_t lds = hv2_lds(op.nup, C, sizeof(T));
    const int grid = (op.ndn + C - 1) / C;
    void (*f)(HvOp, HvArgs);
    if (mode == HV_PLAIN) f = k_hv2<T, ACC, C, HV_PLAIN, KH>;
    else if (mode == HV_ACCUM) f = k_hv2<T, ACC, C, HV_ACCUM, KH>;
    else if (mode == HV_LANCZOS) f = k_hv2<T, ACC, C, HV_LANCZOS, KH>;
    else f = k_hv2<T, double, C, HV_RQ, KH>;
    if (lds > 64 * 1024 &&
        hipFuncSetAttribute((const void *)f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return -1;
    hipLaunchKernelGGL(f, dim3(grid), dim3(kHvBlock), lds, s, op, a);
    if (hipGetLastError() != hipSuccess) return -1;
    return finish_hv(s, mode, grid, a);
}

// Persistent single-orbital H.v with the up-spin hop table resident in LDS.
// One workgroup per CU copies the compressed table (16-bit target rank and
// 8-bit element index per slot, 24 bytes per row) into LDS once; every hop
// record then costs an LDS read instead of an L2 request.  The workgroup
// walks its column groups with the next group's columns loaded into
// registers while the current group is computed (double-buffered LDS
// columns, one barrier per group).  Group order is XCD-aware: the
// workgroups of XCD x (b % 8 == x) share XCD x's contiguous chunk of groups
// and advance through it together, so the down-hop neighbour columns they
// read stay close in the XCD's L2.  Partial sums are per workgroup over its
// groups in a fixed order (deterministic).
constexpr int kHvtSlots = 8;     // stored table slots per row
constexpr int kHvtMaxRows = 4096; // nup bound of the register prefetch

__host__ __device__ constexpr size_t hvt_lds(int nup, int C, size_t esize)
{
    return 2 * align16((size_t)nup * C * esize) + align16((size_t)nup * kHvtSlots * 2) +
           align16((size_t)nup * kHvtSlots);
}

template <typename T, typename ACC, int C, int MODE, int KH, int BS>
__global__ __launch_bounds__(BS) void k_hv1t(HvOp op, HvArgs a)
{
    constexpr int KS = kHvtSlots, NP = kHvtMaxRows / BS;
    if (a.stop && *a.stop) return;
    using Hop = typename HopOf<ACC>::type;
    extern __shared__ __align__(64) unsigned char smem[];
    const int nup = op.nup, ndn = op.ndn;
    const size_t xsz = align16((size_t)nup * C * sizeof(T));
    T *xbuf0 = reinterpret_cast<T *>(smem);
    T *xbuf1 = reinterpret_cast<T *>(smem + xsz);
    uint16_t *tcol = reinterpret_cast<uint16_t *>(smem + 2 * xsz);
    uint8_t *tidx = reinterpret_cast<uint8_t *>(smem + 2 * xsz + align16((size_t)nup * KS * 2));
    __shared__ ACC s_v[256];  // by 8-bit element index: [i] = +V_i, [128 + i] = -V_i, else 0

    for (int i = threadIdx.x; i < 256; i += BS) {
        const int j = i & 127;
        const ACC v = (j < 64 && j < op.nvup) ? (ACC)op.up.vsrc[j] : ACC(0);
        s_v[i] = i < 128 ? v : -v;
    }
    {   // table -> LDS: 16-byte chunks, then the element tail
        const int nc = nup * KS;
        const int c16 = nc * 2 / 16, i16 = nc / 16;
        const uint4 *gc = reinterpret_cast<const uint4 *>(op.up.ccol);
        const uint4 *gi = reinterpret_cast<const uint4 *>(op.up.cidx);
        uint4 *lc = reinterpret_cast<uint4 *>(tcol);
        uint4 *li = reinterpret_cast<uint4 *>(tidx);
        for (int i = threadIdx.x; i < c16; i += BS) lc[i] = gc[i];
        for (int i = threadIdx.x; i < i16; i += BS) li[i] = gi[i];
        for (int i = c16 * 8 + threadIdx.x; i < nc; i += BS) tcol[i] = op.up.ccol[i];
        for (int i = i16 * 16 + threadIdx.x; i < nc; i += BS) tidx[i] = op.up.cidx[i];
    }

    // this workgroup's groups: XCD x's chunk [g, g1), stride = workgroups on x
    const int ngroups = (ndn + C - 1) / C;
    const int nx = gridDim.x < kXcds ? (int)gridDim.x : kXcds;
    const int x = blockIdx.x % nx, wi = blockIdx.x / nx;
    const int wstride = (gridDim.x - x + nx - 1) / nx;
    const int q = ngroups / nx, r = ngroups - q * nx;
    const int gbeg = x * q + (x < r ? x : r);
    const int g1 = gbeg + q + (x < r ? 1 : 0);
    int g = gbeg + wi;

    const T *x_in = static_cast<const T *>(a.x);
    const unsigned char *xb = static_cast<const unsigned char *>(a.x);
    const Hop *dtab = HopOf<ACC>::tab(op.dn);
    T *y = static_cast<T *>(a.y);
    const T *pv = static_cast<const T *>(a.p);

    // column loads of group gg into registers (clamped indices, always in bounds)
    T pf[C][NP];
    auto load_cols = [&](int gg) {
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int idn = gg * C + c < ndn ? gg * C + c : ndn - 1;
            const T *xc = x_in + (long long)idn * nup;
#pragma unroll
            for (int k = 0; k < NP; ++k) {
                const int i = threadIdx.x + k * BS;
                pf[c][k] = xc[i < nup ? i : nup - 1];
            }
        }
    };
    auto store_cols = [&](T *xs) {
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            const int i = threadIdx.x + k * BS;
            if (i < nup) {
                Row<T, C> w;
#pragma unroll
                for (int c = 0; c < C; ++c) w.v[c] = pf[c][k];
                *reinterpret_cast<Row<T, C> *>(xs + i * C) = w;
            }
        }
    };
    if (g < g1) {
        load_cols(g);
        store_cols(xbuf0);
    }
    __syncthreads();

    double sums[3] = {0.0, 0.0, 0.0};
    double sx = 1.0, bsp = 0.0;
    if constexpr (MODE == HV_LANCZOS) {
        if (!lanczos_scales<BS>(a, sx, bsp, blockIdx.x == 0)) return;
    }
    const bool has_p = MODE == HV_LANCZOS && pv != nullptr;

    for (int it = 0; g < g1; ++it) {
        const int gn = g + wstride;
        if (gn < g1) load_cols(gn);  // in flight while group g is computed
        const T *xs = (it & 1) ? xbuf1 : xbuf0;
        const int idn0 = g * C;
        uint32_t coff[C][KH];
        ACC dval[C][KH], ed0[C], edU[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int idn = idn0 + c < ndn ? idn0 + c : ndn - 1;
            const Hop *hr = dtab + (long long)idn * op.dn.K;
#pragma unroll
            for (int k = 0; k < KH; ++k) {
                const Hop h = hr[k];
                coff[c][k] = h.col * (uint32_t)nup * (uint32_t)sizeof(T);
                dval[c][k] = (ACC)h.val;
            }
            const ACC e = energy_of<ACC>(op.dn, idn);
            ed0[c] = e;
            edU[c] = op.dn.mask[idn] ? e + (ACC)op.U : e;
        }
        for (int iup = threadIdx.x; iup < nup; iup += BS) {
            const uint32_t ib = (uint32_t)iup * (uint32_t)sizeof(T);
            T gv[C][KH];
#pragma unroll
            for (int c = 0; c < C; ++c)
#pragma unroll
                for (int k = 0; k < KH; ++k) gv[c][k] = *reinterpret_cast<const T *>(xb + (coff[c][k] + ib));
            const ACC eu = energy_of<ACC>(op.up, iup);
            const bool mu = op.up.mask[iup] != 0;
            T old[C];
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int idn = idn0 + c < ndn ? idn0 + c : ndn - 1;
                const long long e = (long long)idn * nup + iup;
                if constexpr (MODE == HV_ACCUM) old[c] = y[e];
                else if constexpr (MODE == HV_LANCZOS) old[c] = has_p ? pv[e] : T(0);
                else old[c] = T(0);
            }
            const Row<uint16_t, KS> hc = *reinterpret_cast<const Row<uint16_t, KS> *>(tcol + iup * KS);
            const Row<uint8_t, KS> hi = *reinterpret_cast<const Row<uint8_t, KS> *>(tidx + iup * KS);
            const Row<T, C> xr = *reinterpret_cast<const Row<T, C> *>(xs + iup * C);
            ACC acc[C];
#pragma unroll
            for (int c = 0; c < C; ++c) acc[c] = (eu + (mu ? edU[c] : ed0[c])) * (ACC)xr.v[c];
#pragma unroll
            for (int k = 0; k < KH; ++k) {
                const ACC val = s_v[hi.v[k]];
                const Row<T, C> src = *reinterpret_cast<const Row<T, C> *>(xs + (uint32_t)hc.v[k] * C);
#pragma unroll
                for (int c = 0; c < C; ++c) acc[c] += val * (ACC)src.v[c];
            }
#pragma unroll
            for (int k = 0; k < KH; ++k)
#pragma unroll
                for (int c = 0; c < C; ++c) acc[c] += dval[c][k] * (ACC)gv[c][k];
#pragma unroll
            for (int c = 0; c < C; ++c) {
                if (idn0 + c >= ndn) break;
                const long long e = (long long)(idn0 + c) * nup + iup;
                if constexpr (MODE == HV_PLAIN) {
                    y[e] = (T)acc[c];
                } else if constexpr (MODE == HV_ACCUM) {
                    y[e] = (T)((ACC)old[c] + acc[c]);
                } else if constexpr (MODE == HV_LANCZOS) {
                    double rr = sx * (double)acc[c];
                    if (has_p) rr -= bsp * (double)old[c];
                    const T rt = (T)rr;
                    y[e] = rt;
                    sums[0] += (double)xr.v[c] * (double)rt;
                } else {
                    const double xv = (double)xr.v[c], hv = (double)acc[c];
                    sums[0] += xv * hv;
                    sums[1] += hv * hv;
                    sums[2] += xv * xv;
                }
            }
        }
        if (gn < g1) store_cols((it & 1) ? xbuf0 : xbuf1);
        __syncthreads();  // next buffer written; this group's buffer free again
        g = gn;
    }
    if constexpr (MODE == HV_LANCZOS) {
        double mine[1] = {sums[0]};
        block_partial<1, BS>(mine, a.partial);  // alpha_j = s_j * sum (next kernel)
    } else if constexpr (MODE == HV_RQ) {
        block_partial<3, BS>(sums, a.partial);
    }
}

// Usable when the table, two column buffers and the register prefetch fit.
static inline bool hv1t_fits(const HvOp &op, int C, size_t esize, size_t budget)
{
    return op.ncu > 0 && op.up.KC == kHvtSlots && op.nup <= kHvtMaxRows &&
           hvt_lds(op.nup, C, esize) <= budget;
}

template <typename T, typename ACC, int C, int KH, int BS>
static int launch_hv1t_c(hipStream_t s, const HvOp &op, int mode, const HvArgs &a)
{
    const size_t lds = hvt_lds(op.nup, C, sizeof(T));
    const int ngroups = (op.ndn + C - 1) / C;
    const int grid = ngroups < op.ncu ? ngroups : op.ncu;
    void (*f)(HvOp, HvArgs);
    if (mode == HV_PLAIN) f = k_hv1t<T, ACC, C, HV_PLAIN, KH, BS>;
    else if (mode == HV_ACCUM) f = k_hv1t<T, ACC, C, HV_ACCUM, KH, BS>;
    else if (mode == HV_LANCZOS) f = k_hv1t<T, ACC, C, HV_LANCZOS, KH, BS>;
    else f = k_hv1t<T, double, C, HV_RQ, KH, BS>;
    if (hipFuncSetAttribute((const void *)f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
        hipSuccess)
        return -1;
    hipLaunchKernelGGL(f, dim3(grid), dim3(BS), lds, s, op, a);
    if (hipGetLastError() != hipSuccess) return -1;
    return finish_hv(s, mode, grid, a);
}

template <typename T, typename ACC, int C, int KH>
static int launch_hv1_c(hipStream_t s, const HvOp &op, int mode, const HvArgs &a)
{
    const size_t lds = (size_t)op.nup * C * sizeof(T);
    // one workgroup per column group: the workgroup scheduler then overlaps
    // one group's column staging with other groups' compute (measured faster
    // on MI355X than a grid-stride loop with a few workgroups per CU)
    const int grid = (op.ndn + C - 1) / C;
    void (*f)(HvOp, HvArgs);
    if (mode == HV_PLAIN) f = k_hv1<T, ACC, C, HV_PLAIN, KH>;
    else if (mode == HV_ACCUM) f = k_hv1<T, ACC, C, HV_ACCUM, KH>;
    else if (mode == HV_LANCZOS) f = k_hv1<T, ACC, C, HV_LANCZOS, KH>;
    else if (mode == HV_FUSED) f = k_hv1<T, ACC, C, HV_FUSED, KH>;
    else f = k_hv1<T, double, C, HV_RQ, KH>;
    if (lds > 64 * 1024) {
        if (hipFuncSetAttribute((const void *)f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds) != hipSuccess)
            return -1;
    }
    hipLaunchKernelGGL(f, dim3(grid), dim3(kHvBlock), lds, s, op, a);
    if (hipGetLastError() != hipSuccess) return -1;
    return finish_hv(s, mode, grid, a);
}

template <typename T, typename ACC, int KH>
static int launch_hv1_k(hipStream_t s, const HvOp &op, int mode, int cols, const HvArgs &a)
{
    switch (cols) {
    case 1: return launch_hv1_c<T, ACC, 1, KH>(s, op, mode, a);
    case 2: return launch_hv1_c<T, ACC, 2, KH>(s, op, mode, a);
    case 4:
    case 8: return launch_hv1_c<T, ACC, 4, KH>(s, op, mode, a);  // 8 columns would spill
    default: return -2;
    }
}

// Dynamic LDS budget of one workgroup: gfx950 has 160 KiB per CU; keep 8 KiB
// for the static per-column hop lists and reduction scratch.
constexpr size_t kLdsBudget = 152 * 1024;

template <typename T, typename ACC, int C, bool KAN, int KH>
static int launch_hv_c(hipStream_t s, const HvOp &op, int mode, const HvArgs &a)
{
    const size_t lds = (size_t)op.nup * C * sizeof(T);
    const int grid = (op.ndn + C - 1) / C;
    void (*f)(HvOp, HvArgs);
    if (mode == HV_PLAIN) f = k_hv<T, ACC, C, HV_PLAIN, KAN, KH>;
    else if (mode == HV_ACCUM) f = k_hv<T, ACC, C, HV_ACCUM, KAN, KH>;
    else if (mode == HV_LANCZOS) f = k_hv<T, ACC, C, HV_LANCZOS, KAN, KH>;
    else f = k_hv<T, double, C, HV_RQ, KAN, KH>;
    if (lds > 64 * 1024) {
        if (hipFuncSetAttribute((const void *)f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds) != hipSuccess)
            return -1;
    }
    hipLaunchKernelGGL(f, dim3(grid), dim3(kHvBlock), lds, s, op, a);
    if (hipGetLastError() != hipSuccess) return -1;
    return finish_hv(s, mode, grid, a);
}

template <typename T, typename ACC, bool KAN, int KH>
static int launch_hv_k(hipStream_t s, const HvOp &op, int mode, int cols, const HvArgs &a)
{
    switch (cols) {
    case 1: return launch_hv_c<T, ACC, 1, KAN, KH>(s, op, mode, a);
    case 2: return launch_hv_c<T, ACC, 2, KAN, KH>(s, op, mode, a);
    case 4: return launch_hv_c<T, ACC, 4, KAN, KH>(s, op, mode, a);
    case 8: return launch_hv_c<T, ACC, 8, KAN, KH>(s, op, mode, a);
    default: return -2;
    }
}

// Two columns per workgroup when the tuning asks for more than one and their
// LDS fits.  Measured on MI355X at Ns=14, sector (7,7), FP32 vectors: the
// plain H.v takes 79 us with two columns against 100 us with one, the solve
// 135 against 141 ms (the LDS of two columns leaves room for two workgroups
// per CU, that of one column for four).  FP64: one column is faster (262
// against 280 ms per solve).  Contexts default accordingly.
template <typename T, typename ACC, int KH>
static int launch_hv2_k(hipStream_t s, const HvOp &op, int mode, int cols, const HvArgs &a)
{
    if (cols >= 2 && hv2_lds(op.nup, 2, sizeof(T)) <= kLdsBudget) return launch_hv2_c<T, ACC, 2, KH>(s, op, mode, a);
    return launch_hv2_c<T, ACC, 1, KH>(s, op, mode, a);
}

template <typename T, typename ACC>
static int launch_hv_a(hipStream_t s, const HvOp &op, int mode, int cols, const HvArgs &a)
{
    const bool kan = op.norb >= 2;
    const int km = op.up.kmax > op.dn.kmax ? op.up.kmax : op.dn.kmax;
    if (mode == HV_FUSED) {  // the single-orbital tile only
        if (op.norb != 1 || op.sched >= 16) return -4;
        if (op.up.KC == 8 && km <= 7) return launch_hv1_k<T, ACC, 7>(s, op, mode, cols, a);
        if (op.up.KC == 8 && km <= 8) return launch_hv1_k<T, ACC, 8>(s, op, mode, cols, a);
        if (op.up.KC == 16 && km <= 16) return launch_hv1_k<T, ACC, 16>(s, op, mode, cols, a);
        return -4;
    }
    // two orbitals: the specialized kernel (tables with <= 16 slots, LDS for a column and its target)
    if (op.norb == kNorb2 && op.up.KC >= 8 && op.up.KC == op.dn.KC && km <= 16 && op.up.hrec &&
        hv2_lds(op.nup, 1, sizeof(T)) <= kLdsBudget && op.sched < 16) {
        if (km <= 8 && op.up.KC == 8) return launch_hv2_k<T, ACC, 8>(s, op, mode, cols, a);
        if (op.up.KC == 16) {
            if (km <= 12) return launch_hv2_k<T, ACC, 12>(s, op, mode, cols, a);
            return launch_hv2_k<T, ACC, 16>(s, op, mode, cols, a);
        }
    }
    if (kan) return launch_hv_k<T, ACC, true, 0>(s, op, mode, cols, a);
    // persistent table-in-LDS kernel: schedule 16 = 1 column per group and
    // 512 threads, 32 = 1 column and 1024 threads, 64 = 2 columns and 1024
    // threads (1 column when 2 do not fit in LDS)
    if (op.norb == 1 && km <= 8 && op.sched >= 16) {
        int pc = op.sched >= 64 ? 2 : 1;
        if (!hv1t_fits(op, pc, sizeof(T), kLdsBudget)) pc = 1;
        if (hv1t_fits(op, pc, sizeof(T), kLdsBudget)) {
            if (pc == 2) {
                return km <= 7 ? launch_hv1t_c<T, ACC, 2, 7, 1024>(s, op, mode, a)
                               : launch_hv1t_c<T, ACC, 2, 8, 1024>(s, op, mode, a);
            }
            if (op.sched >= 32) {
                return km <= 7 ? launch_hv1t_c<T, ACC, 1, 7, 1024>(s, op, mode, a)
                               : launch_hv1t_c<T, ACC, 1, 8, 1024>(s, op, mode, a);
            }
            return km <= 7 ? launch_hv1t_c<T, ACC, 1, 7, 512>(s, op, mode, a)
                           : launch_hv1t_c<T, ACC, 1, 8, 512>(s, op, mode, a);
        }
    }
    if (op.norb == 1 && op.up.KC == 8 && km <= 7) return launch_hv1_k<T, ACC, 7>(s, op, mode, cols, a);
    if (op.norb == 1 && op.up.KC == 8 && km <= 8) return launch_hv1_k<T, ACC, 8>(s, op, mode, cols, a);
    if (op.norb == 1 && op.up.KC == 16 && km <= 16) return launch_hv1_k<T, ACC, 16>(s, op, mode, cols, a);
    if (km <= 8) return launch_hv_k<T, ACC, false, 8>(s, op, mode, cols, a);
    if (km <= 16) return launch_hv_k<T, ACC, false, 16>(s, op, mode, cols, a);
    return launch_hv_k<T, ACC, false, 0>(s, op, mode, cols, a);
}

template <typename T>
int launch_hv(hipStream_t s, const HvOp &op_in, int mode, int cols, bool acc64, const HvArgs &a)
{
    // cols 1/2/4/8: one workgroup per column group; 16/32/64: persistent
    // workgroups with 1/2/4 columns (falls back to 1 column per workgroup)
    HvOp op = op_in;
    if (mode == HV_PLAIN && a.accumulate) mode = HV_ACCUM;
    op.sched = cols >= 16 ? cols : 0;
    if (cols >= 16) cols = 1;
    while (cols > 1 && (size_t)op.nup * cols * sizeof(T) > kLdsBudget) cols >>= 1;
    if ((size_t)op.nup * cols * sizeof(T) > kLdsBudget) return -3;
    if constexpr (sizeof(T) == 8) {
        return launch_hv_a<T, double>(s, op, mode, cols, a);
    } else {
        if (acc64 || mode == HV_RQ) return launch_hv_a<T, double>(s, op, mode, cols, a);
        return launch_hv_a<T, float>(s, op, mode, cols, a);
    }
}
template int launch_hv<float>(hipStream_t, const HvOp &, int, int, bool, const HvArgs &);
template int launch_hv<double>(hipStream_t, const HvOp &, int, int, bool, const HvArgs &);

bool hv_fused_supported(const HvOp &op, int cols, size_t esize)
{
    const int km = op.up.kmax > op.dn.kmax ? op.up.kmax : op.dn.kmax;
    return op.norb == 1 && cols < 16 && ((op.up.KC == 8 && km <= 8) || (op.up.KC == 16 && km <= 16)) &&
           (size_t)op.nup * esize <= kLdsBudget;
}

// ----------------------------------------------------------- vector ops

static inline int grid_for(long long n)
{
    long long g = (n + (long long)kBlock * 8 - 1) / ((long long)kBlock * 8);
    if (g < 1) g = 1;
    if (g > 2048) g = 2048;
    return (int)g;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_norm(const T *x, long long n, double *partial, const int *stop)
{
    if (stop && *stop) return;
    double v[1] = {0.0};
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (long long)gridDim.x * kBlock) {
        const double t = (double)x[i];
        v[0] += t * t;
    }
    block_partial<1>(v, partial);
}

template <typename T>
int launch_norm(hipStream_t s, const T *x, long long n, double *partial,
                double *norm_out, double *scale_out, const int *stop)
{
    const int grid = grid_for(n);
    hipLaunchKernelGGL(k_norm<T>, dim3(grid), dim3(kBlock), 0, s, x, n, partial, stop);
    if (hipGetLastError() != hipSuccess) return -1;
    return launch_finish<1>(s, partial, FinArgs{FIN_NORM, grid, nullptr, norm_out, scale_out,
                                                nullptr, nullptr, 0, 0, nullptr, nullptr});
}
template int launch_norm<float>(hipStream_t, const float *, long long, double *,
                                double *, double *, const int *);
template int launch_norm<double>(hipStream_t, const double *, long long, double *,
                                 double *, double *, const int *);

// Lanczos step j, second half: alpha_j = s_j * sum(hv_part) (every
// workgroup reduces the H.v partials in the same order; workgroup 0 records
// alpha_j), r <- r - alpha_j s_j u_j, partial sums of |r|^2 for beta_{j+1}.
// Few large workgroups: every workgroup of the next H.v sums this kernel's
// partials before its row loop, and every workgroup here sums the H.v's, so
// the partial counts are kept small (1024 x 256 threads as many as 4096 x
// 64 would be, a quarter of the partials of 256-thread workgroups).
constexpr int kAxpyBlock = 1024;
constexpr int kAxpyGrid = kAxpyMaxGrid;

template <typename T, int BS>
__global__ __launch_bounds__(BS) void k_axpy_norm(T *r, const T *u, long long n,
                                                      const double *hv_part, int hv_n,
                                                      const double *scale, double *alpha,
                                                      int step, double *partial, const int *stop)
{
    if (*stop) return;
    const double sj = scale[step];
    const double al = sj * sum_partials<BS>(hv_part, hv_n);
    if (blockIdx.x == 0 && threadIdx.x == 0) alpha[step] = al;
    const double coef = al * sj;
    // 16-byte vectors (vectors are 256-byte aligned); workgroup 0 also takes
    // the n % V tail.  The update itself is FP64.
    constexpr int V = 16 / sizeof(T);
    using Vec = Row<T, V>;
    const long long nv = n / V, stride = (long long)gridDim.x * BS;
    Vec *rv = reinterpret_cast<Vec *>(r);
    const Vec *uv = reinterpret_cast<const Vec *>(u);
    double v[1] = {0.0};
    for (long long q = (long long)blockIdx.x * BS + threadIdx.x; q < nv; q += stride) {
        const Vec a = rv[q], b = uv[q];
        Vec w;
#pragma unroll
        for (int i = 0; i < V; ++i) {
            w.v[i] = (T)((double)a.v[i] - coef * (double)b.v[i]);
            v[0] += (double)w.v[i] * (double)w.v[i];
        }
        rv[q] = w;
    }
    if (blockIdx.x == 0)
        for (long long i = nv * V + threadIdx.x; i < n; i += BS) {
            const T t = (T)((double)r[i] - coef * (double)u[i]);
            r[i] = t;
            v[0] += (double)t * (double)t;
        }
    block_partial<1, BS>(v, partial);
}

template <typename T>
int launch_axpy_norm(hipStream_t s, T *r, const T *u, long long n, const double *hv_part,
                     int hv_n, const double *scale, double *alpha, int step, double *partial,
                     int *stop)
{
    if (((uintptr_t)r | (uintptr_t)u) & 15) return -1;  // vector kernels need 16-byte alignment
    long long g = (n + (long long)kAxpyBlock * 16 - 1) / ((long long)kAxpyBlock * 16);
    const int grid = g < 1 ? 1 : (g > kAxpyGrid ? kAxpyGrid : (int)g);
    hipLaunchKernelGGL((k_axpy_norm<T, kAxpyBlock>), dim3(grid), dim3(kAxpyBlock), 0, s, r, u, n, hv_part,
                       hv_n, scale, alpha, step, partial, stop);
    return hipGetLastError() == hipSuccess ? grid : -1;
}
template int launch_axpy_norm<float>(hipStream_t, float *, const float *, long long,
                                     const double *, int, const double *, double *, int,
                                     double *, int *);
template int launch_axpy_norm<double>(hipStream_t, double *, const double *, long long,
                                      const double *, int, const double *, double *, int,
                                      double *, int *);

// Counter-based U[0,1): splitmix64(seed * K + i) >> 11 scaled by 2^-53.
__device__ __forceinline__ uint64_t splitmix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Streaming reference passes (edg_time_stream): what plain 16-byte loads and
// stores reach on this device, the yardstick of the H.v's fabric traffic.
// One 16-byte element per thread, the grid covering the vector once.
// Measured on MI355X at 47 MB per vector: a grid-stride loop over a capped
// grid (part of the grid takes a second pass) and four elements per thread
// with the loads first were both slower (15.4 and 23.7 against 14.2 us per
// copy in a kernel trace).
__global__ __launch_bounds__(kBlock) void k_stream_copy(const float4 *x, float4 *y, long long n4)
{
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n4) y[i] = x[i];
}
__global__ __launch_bounds__(kBlock) void k_stream_add(const float4 *x, const float4 *z, float4 *y, long long n4)
{
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n4) {
        const float4 a = x[i], b = z[i];
        y[i] = make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
    }
}
static int stream_grid(long long n4) { return (int)((n4 + kBlock - 1) / kBlock); }
int launch_stream_copy(hipStream_t s, const float *x, float *y, long long n)
{
    const long long n4 = n / 4;
    hipLaunchKernelGGL(k_stream_copy, dim3(stream_grid(n4)), dim3(kBlock), 0, s, (const float4 *)x, (float4 *)y, n4);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
int launch_stream_add(hipStream_t s, const float *x, const float *z, float *y, long long n)
{
    const long long n4 = n / 4;
    hipLaunchKernelGGL(k_stream_add, dim3(stream_grid(n4)), dim3(kBlock), 0, s, (const float4 *)x,
                       (const float4 *)z, (float4 *)y, n4);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// partial non-null: also the partial sums of |x|^2, in k_norm's element
// order (same grid), so the norm is the one a separate k_norm would give.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_start(T *x, long long n, unsigned long long seed, double *partial)
{
    double v[1] = {0.0};
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (long long)gridDim.x * kBlock) {
        const T t = (T)((double)(splitmix64(seed * 0xD1B54A32D192ED03ull + (uint64_t)i) >> 11) *
                        0x1.0p-53);
        x[i] = t;
        v[0] += (double)t * (double)t;
    }
    if (partial) block_partial<1>(v, partial);
}

template <typename T>
int launch_start(hipStream_t s, T *x, long long n, unsigned long long seed, double *partial,
                 double *norm_out, double *scale_out)
{
    const int grid = grid_for(n);
    hipLaunchKernelGGL(k_start<T>, dim3(grid), dim3(kBlock), 0, s, x, n, seed, partial);
    if (hipGetLastError() != hipSuccess) return -1;
    if (!partial) return 0;
    return launch_finish<1>(s, partial, FinArgs{FIN_NORM, grid, nullptr, norm_out, scale_out,
                                                nullptr, nullptr, 0, 0, nullptr, nullptr});
}
template int launch_start<float>(hipStream_t, float *, long long, unsigned long long, double *, double *,
                                 double *);
template int launch_start<double>(hipStream_t, double *, long long, unsigned long long, double *, double *,
                                  double *);

// out = sum_j coef_j * basis_j (FP64 accumulation, in j order), plus
// |out|^2.  Each thread issues the loads of eight vectors before their
// FMAs: with one load per step (nvec = 200) a small sector's combine waited
// on 200 memory round trips in a row.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_combine(T *out, const T *basis, long long stride,
                                                    int nvec, const double *coef, long long n,
                                                    double *partial)
{
    __shared__ double c[1024];
    for (int j = threadIdx.x; j < nvec && j < 1024; j += kBlock) c[j] = coef[j];
    __syncthreads();
    double v[1] = {0.0};
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (long long)gridDim.x * kBlock) {
        double acc = 0.0;
        int j = 0;
        for (; j + 8 <= nvec; j += 8) {
            T b[8];
#pragma unroll
            for (int r = 0; r < 8; ++r) b[r] = basis[(long long)(j + r) * stride + i];
#pragma unroll
            for (int r = 0; r < 8; ++r) acc += c[j + r] * (double)b[r];
        }
        for (; j < nvec; ++j) acc += c[j] * (double)basis[(long long)j * stride + i];
        const T t = (T)acc;
        out[i] = t;
        v[0] += (double)t * (double)t;
    }
    block_partial<1>(v, partial);
}

template <typename T>
int launch_combine(hipStream_t s, T *out, const T *basis, long long stride, int nvec,
                   const double *coef, long long n, double *partial,
                   double *norm_out, double *scale_out)
{
    if (nvec > 1024) return -2;
    // one element per thread up to 4096 workgroups: every element is a sum
    // over nvec vectors, so small sectors need the threads
    const long long g = (n + kBlock - 1) / kBlock;
    const int grid = (int)(g < 1 ? 1 : (g > 4096 ? 4096 : g));
    hipLaunchKernelGGL(k_combine<T>, dim3(grid), dim3(kBlock), 0, s, out, basis, stride,
                       nvec, coef, n, partial);
    if (hipGetLastError() != hipSuccess) return -1;
    return launch_finish<1>(s, partial, FinArgs{FIN_NORM, grid, nullptr, norm_out, scale_out,
                                                nullptr, nullptr, 0, 0, nullptr, nullptr});
}
template int launch_combine<float>(hipStream_t, float *, const float *, long long, int,
                                   const double *, long long, double *, double *,
                                   double *);
template int launch_combine<double>(hipStream_t, double *, const double *, long long, int,
                                    const double *, long long, double *, double *,
                                    double *);

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scale(T *x, long long n, const double *scale)
{
    const double sc = *scale;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (long long)gridDim.x * kBlock)
        x[i] = (T)((double)x[i] * sc);
}
template <typename T>
int launch_scale(hipStream_t s, T *x, long long n, const double *scale)
{
    hipLaunchKernelGGL(k_scale<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, x, n, scale);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
template int launch_scale<float>(hipStream_t, float *, long long, const double *);
template int launch_scale<double>(hipStream_t, double *, long long, const double *);

template <typename T>
__global__ __launch_bounds__(kBlock) void k_from_f64(T *dst, const double *src, long long n)
{
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (long long)gridDim.x * kBlock)
        dst[i] = (T)src[i];
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_to_f64(double *dst, const T *src, long long n)
{
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (long long)gridDim.x * kBlock)
        dst[i] = (double)src[i];
}
template <typename T>
int launch_from_f64(hipStream_t s, T *dst, const double *src, long long n)
{
    hipLaunchKernelGGL(k_from_f64<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, dst, src, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
template <typename T>
int launch_to_f64(hipStream_t s, double *dst, const T *src, long long n)
{
    hipLaunchKernelGGL(k_to_f64<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, dst, src, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
template int launch_from_f64<float>(hipStream_t, float *, const double *, long long);
template int launch_from_f64<double>(hipStream_t, double *, const double *, long long);
template int launch_to_f64<float>(hipStream_t, double *, const float *, long long);
template int launch_to_f64<double>(hipStream_t, double *, const double *, long long);

// ------------------------------------------------------------ apply_c

// map[r'] for output configs r' of the changed spin (ed_operator.f90:31-54):
// c+ (pm=1) needs bit set in the output pattern, source = bit cleared;
// c (pm=2) needs bit clear, source = bit set.  Sign = parity of the output
// pattern's bits below `bit` (same-spin part of the reference's sgn).
__global__ __launch_bounds__(kBlock) void k_cmap(const uint32_t *pattern_out, int n_out,
                                                 int nsite, int ne_in, int bit, int pm,
                                                 uint32_t *map)
{
    __shared__ uint32_t B[16][16];
    for (int e = threadIdx.x; e < 256; e += kBlock) {
        int n = e >> 4, k = e & 15;
        uint32_t c = 0;
        if (k <= n) {
            c = 1;
            for (int i = 1; i <= k; ++i) c = c * (uint32_t)(n - k + i) / (uint32_t)i;
        }
        B[n][k] = c;
    }
    __syncthreads();
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n_out) return;
    const uint32_t p = pattern_out[r];
    const uint32_t set = (p >> bit) & 1u;
    if ((pm == 1 && !set) || (pm == 2 && set)) { map[r] = kInvalid; return; }
    const uint32_t q = p ^ (1u << bit);
    uint32_t s = 0;
    int k = 0;
    for (int b = 0; b < nsite; ++b)
        if ((q >> b) & 1u) { ++k; s += binom_tab(B, b, k); }
    (void)ne_in;
    const uint32_t odd = __popc(p & ((1u << bit) - 1u)) & 1u;
    map[r] = s | (odd << 31);
}

int launch_cmap(hipStream_t s, const uint32_t *pattern_out, int n_out, int nsite, int ne_in,
                int bit, int pm, uint32_t *map)
{
    hipLaunchKernelGGL(k_cmap, dim3((n_out + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                       pattern_out, n_out, nsite, ne_in, bit, pm, map);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// spin 1: out(iup', idn) = sgn * in(map(iup'), idn)
// spin 2: out(iup, idn') = sgn * (-1)^{N_up} * in(iup, map(idn'))
// partial non-null: also the partial sums of |out|^2 (k_norm's order, as
// k_start).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_apply_c(const T *in, T *out, int nup_out,
                                                    int ndn_out, int nup_in, int spin,
                                                    const uint32_t *map, int global_odd,
                                                    double *partial)
{
    const long long n = (long long)nup_out * ndn_out;
    double acc[1] = {0.0};
    for (long long e = (long long)blockIdx.x * kBlock + threadIdx.x; e < n;
         e += (long long)gridDim.x * kBlock) {
        const int idn = (int)(e / nup_out);
        const int iup = (int)(e - (long long)idn * nup_out);
        T v = T(0);
        if (spin == 1) {
            const uint32_t m = map[iup];
            if (m != kInvalid) {
                const T t = in[(long long)idn * nup_in + (m & 0x7FFFFFFFu)];
                v = (m >> 31) ? -t : t;
            }
        } else {
            const uint32_t m = map[idn];
            if (m != kInvalid) {
                const T t = in[(long long)(m & 0x7FFFFFFFu) * nup_in + iup];
                v = (((m >> 31) ^ (uint32_t)global_odd) & 1u) ? -t : t;
            }
        }
        out[e] = v;
        acc[0] += (double)v * (double)v;
    }
    if (partial) block_partial<1>(acc, partial);
}
template <typename T>
int launch_apply_c(hipStream_t s, const T *in, T *out, int nup_out, int ndn_out, int nup_in,
                   int spin, const uint32_t *map, int global_odd, double *partial, double *norm_out,
                   double *scale_out)
{
    const long long n = (long long)nup_out * ndn_out;
    const int grid = grid_for(n);
    hipLaunchKernelGGL(k_apply_c<T>, dim3(grid), dim3(kBlock), 0, s, in, out, nup_out,
                       ndn_out, nup_in, spin, map, global_odd, partial);
    if (hipGetLastError() != hipSuccess) return -1;
    if (!partial) return 0;
    return launch_finish<1>(s, partial, FinArgs{FIN_NORM, grid, nullptr, norm_out, scale_out,
                                                nullptr, nullptr, 0, 0, nullptr, nullptr});
}
template int launch_apply_c<float>(hipStream_t, const float *, float *, int, int, int, int,
                                   const uint32_t *, int, double *, double *, double *);
template int launch_apply_c<double>(hipStream_t, const double *, double *, int, int, int, int,
                                    const uint32_t *, int, double *, double *, double *);

// -------------------------------------------------------- observables

// post_process.f90:63-112: the per-workgroup partials of n(m,up), n(m,dn)
// and docc(m), with gs scaled by *scale; the finisher writes out[m] =
// n(m,up), out[norb+m] = n(m,dn), out[2 norb + m] = docc(m).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_observables(const T *gs, HvOp op,
                                                        const double *scale, double *partial)
{
    constexpr int NS = 3 * kMaxNorb;
    double v[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) v[s] = 0.0;
    const double sc = scale ? *scale : 1.0;
    const long long n = (long long)op.nup * op.ndn;
    const int norb = op.norb;
    for (long long e = (long long)blockIdx.x * kBlock + threadIdx.x; e < n;
         e += (long long)gridDim.x * kBlock) {
        const int idn = (int)(e / op.nup);
        const int iup = (int)(e - (long long)idn * op.nup);
        const double g = (double)gs[e] * sc;
        const double w = g * g;
        const int mu = op.up.mask[iup], md = op.dn.mask[idn];
#pragma unroll
        for (int m = 0; m < kMaxNorb; ++m) {
            if (m < norb) {
                if ((mu >> m) & 1) v[m] += w;
                if ((md >> m) & 1) v[kMaxNorb + m] += w;
                if ((mu >> m) & (md >> m) & 1) v[2 * kMaxNorb + m] += w;
            }
        }
    }
    block_partial<NS>(v, partial);
}
template <typename T>
int launch_observables(hipStream_t s, const T *gs, const HvOp &op, const double *scale,
                       double *partial, double *out)
{
    const long long n = (long long)op.nup * op.ndn;
    const int grid = grid_for(n);
    hipLaunchKernelGGL(k_observables<T>, dim3(grid), dim3(kBlock), 0, s, gs, op, scale,
                       partial);
    if (hipGetLastError() != hipSuccess) return -1;
    return launch_finish<3 * kMaxNorb>(s, partial, FinArgs{FIN_OBS, grid, nullptr, out, nullptr,
                                                           nullptr, nullptr, 0, op.norb, nullptr, nullptr});
}
template int launch_observables<float>(hipStream_t, const float *, const HvOp &, const double *,
                                       double *, double *);
template int launch_observables<double>(hipStream_t, const double *, const HvOp &,
                                        const double *, double *, double *);

// --------------------------------------------------- continued fraction

struct cplx {
    double re, im;
};
__device__ __forceinline__ cplx cdiv(double num, cplx d)
{
    // num / d for real num, Smith's scaling, without a branch: the lanes of
    // a wave take either ordering (a branch ran both sides) and the level's
    // two divisions are the only ones on the recursion's critical path
    const bool big_re = fabs(d.re) >= fabs(d.im);
    const double p = big_re ? d.re : d.im, q = big_re ? d.im : d.re;
    const double r = q / p, inv = 1.0 / (p + q * r);
    const double a = num * inv, b = num * r * inv;
    return big_re ? cplx{a, -b} : cplx{b, -a};
}

// numeric_utils.f90:11-44: f = b1^2/(z -+ a1 - b2^2/(z -+ a2 - ...)),
// evaluated bottom-up; the last level is applied after the loop even for n=1.
// a and b2 = b*b are read from LDS.
__device__ cplx cfrac(cplx z, int n, const double *a, const double *b2, double sgn)
{
    cplx f = cdiv(b2[n - 1], {z.re + sgn * a[n - 1], z.im});
    for (int i = n - 2; i >= 1; --i) f = cdiv(b2[i], {z.re + sgn * a[i] - f.re, z.im - f.im});
    return cdiv(b2[0], {z.re + sgn * a[0] - f.re, z.im - f.im});
}

// g[iw] (+)= CF_p(E0 + i w; ap, bp) + CF_m(-E0 + i w; am, bm).  A workgroup
// of 128 threads takes 64 frequencies: threads 0-63 evaluate CF_p and 64-127
// CF_m of the same frequencies, and the pair adds in LDS.  The recursion is
// a serial chain of divisions per thread, so the coefficients (and b^2) are
// staged in LDS first: read from global memory level by level, every level
// would wait on an L2 round trip.
constexpr int kCfPer = 64;  // frequencies per workgroup
__global__ __launch_bounds__(2 * kCfPer) void k_cf(int nw, const double *omega, double e0, int np, const double *ap,
                                                   const double *bp, int nm, const double *am, const double *bm,
                                                   double *g, int accumulate)
{
    extern __shared__ double cfs[];  // ap[np] | bp^2[np] | am[nm] | bm^2[nm]
    __shared__ cplx s_m[kCfPer];
    double *sap = cfs, *sbp = cfs + np, *sam = cfs + 2 * np, *sbm = cfs + 2 * np + nm;
    for (int i = threadIdx.x; i < np; i += blockDim.x) {
        sap[i] = ap[i];
        sbp[i] = bp[i] * bp[i];
    }
    for (int i = threadIdx.x; i < nm; i += blockDim.x) {
        sam[i] = am[i];
        sbm[i] = bm[i] * bm[i];
    }
    __syncthreads();
    const int half = threadIdx.x / kCfPer, k = threadIdx.x % kCfPer;
    const int iw = blockIdx.x * kCfPer + k;
    cplx f{0.0, 0.0};
    if (iw < nw) {
        if (half == 0 && np > 0) f = cfrac({e0, omega[iw]}, np, sap, sbp, -1.0);
        if (half == 1 && nm > 0) f = cfrac({-e0, omega[iw]}, nm, sam, sbm, 1.0);
    }
    if (half == 1) s_m[k] = f;
    __syncthreads();
    if (half == 1 || iw >= nw) return;
    const double re = f.re + s_m[k].re, im = f.im + s_m[k].im;
    if (accumulate) {
        g[2 * iw] += re;
        g[2 * iw + 1] += im;
    } else {
        g[2 * iw] = re;
        g[2 * iw + 1] = im;
    }
}

// Self-energy of post_process.f90 dump_sigma (:121-185) and the
// quasiparticle weight of :40-58, one thread per (w_n, orbital, spin):
//   D(i w) = sum_l vmk(m,l,s)^2 / (i w - ek(l,s))   (dmft_green.f90:24-45)
//   Sigma  = i w + mu - em(m,s) - D - 1/G_cl
//   Z(m,s) = 1 / (1 - Im Sigma(w_1) / w_1)
// params = em(norb,2) | ek(nbath,2) | vmk(norb,nbath,2), column-major;
// G and Sigma interleaved complex (nw, norb, nspin), column-major.
__global__ __launch_bounds__(kBlock) void k_sigma(int nw, int norb, int nbath, int nspin, const double *omega,
                                                 const double *params, double mu, const double *G,
                                                 double *Sigma, double *Z)
{
    const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (t >= (long long)nw * norb * nspin) return;
    const int iw = (int)(t % nw), m = (int)((t / nw) % norb), s = (int)(t / ((long long)nw * norb));
    const double *em = params, *ek = params + 2 * norb, *vmk = params + 2 * norb + 2 * nbath;
    const double w = omega[iw];
    double dr = 0.0, di = 0.0;
    for (int l = 0; l < nbath; ++l) {
        const double v = vmk[(size_t)s * nbath * norb + (size_t)l * norb + m];
        const double e = ek[(size_t)s * nbath + l];
        const double q = v * v / (w * w + e * e);  // v^2 / (i w - e) = v^2 (-e - i w) / (w^2 + e^2)
        dr -= e * q;
        di -= w * q;
    }
    const double gr = G[2 * t], gi = G[2 * t + 1];
    const double g2 = gr * gr + gi * gi;
    const double sr = mu - em[(size_t)s * norb + m] - dr - gr / g2;
    const double si = w - di + gi / g2;
    Sigma[2 * t] = sr;
    Sigma[2 * t + 1] = si;
    if (iw == 0) Z[(size_t)s * norb + m] = 1.0 / (1.0 - si / w);
}

int launch_sigma(hipStream_t s, int nw, int norb, int nbath, int nspin, const double *omega,
                 const double *params, double mu, const double *G, double *Sigma, double *Z)
{
    const long long n = (long long)nw * norb * nspin;
    hipLaunchKernelGGL(k_sigma, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, nw, norb, nbath,
                       nspin, omega, params, mu, G, Sigma, Z);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_cf(hipStream_t s, int nw, const double *omega, double e0, int np, const double *ap,
              const double *bp, int nm, const double *am, const double *bm, double *g,
              int accumulate)
{
    if (np < 0 || nm < 0 || np > 1000 || nm > 1000) return -1;  // chains are at most 1000 levels
    const size_t lds = sizeof(double) * 2 * ((size_t)np + nm);
    hipLaunchKernelGGL(k_cf, dim3((nw + kCfPer - 1) / kCfPer), dim3(2 * kCfPer), lds, s, nw, omega, e0, np, ap, bp,
                       nm, am, bm, g, accumulate);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace edg
