// Single-orbital H.v tile shared by the single-chain kernel (kernels.hip)
// and the batched long-column kernel (batch.hip).
#pragma once
#include "device_util.h"

namespace edg {

// Single-orbital H.v tile (norb == 1) over the sector columns
// [idn0, idn0 + C), partial sums written at partial slot `slot`; `lead`
// marks the workgroup that records beta_j and s_j.  The diagonal is
// E_up(iup) + E_dn(idn) + U n_up n_dn with n the impurity bit, and both spins
// hop through tables whose unused slots are self references with a zero
// element.
//   up spin:   hrec[iup][KS] records (KS = 8 or 16 stored slots, the first KH
//              processed), rank | (value index << 16), gathered from the C
//              columns staged in LDS;
//   down spin: the C columns' neighbour-column byte offsets and elements are
//              wave-uniform, read once into scalar registers / LDS, and each
//              hop is one coalesced load of a neighbour column.
// The column staging issues all its loads before the LDS stores.  The row
// loop is written for U rows per thread in flight; U = 1 measured fastest
// on MI355X (Ns=14: 46 us against 50 us for U = 2 and 61 us for U = 4 at two
// columns per workgroup): the kernel is bound by the L1/L2 request stream
// of the neighbour-column loads, not by their latency, and fewer registers
// keep more waves resident.
template <typename T, typename ACC, int C, int MODE, int KH>
__device__ __forceinline__ void hv1_tile(const HvOp &op, const HvArgs &a, int idn0, int slot, bool lead)
{
    constexpr int KS = KH <= 8 ? 8 : 16;
    constexpr int U = 1;
    constexpr int kStage = C <= 8 ? 8 / C : 1;
    using Hop = typename HopOf<ACC>::type;
    extern __shared__ __align__(64) unsigned char smem[];
    T *xs = reinterpret_cast<T *>(smem);
    __shared__ Row<ACC, C> s_dval[KH];      // down-hop elements, slot-major
    __shared__ ACC s_v[2 * (kMaxHops + 1)];  // up-spin elements: [idx] = +V, [65 + idx] = -V, idx 64 = 0

    const int nup = op.nup, ndn = op.ndn;
    const T *x = static_cast<const T *>(a.x);
    const unsigned char *xb = static_cast<const unsigned char *>(a.x);
    const Hop *dtab = HopOf<ACC>::tab(op.dn);
    for (int i = threadIdx.x; i <= kMaxHops; i += kHvBlock) {
        const ACC v = i < op.nvup ? (ACC)gload(op.up.vsrc + i) : ACC(0);
        s_v[i] = v;
        s_v[kMaxHops + 1 + i] = -v;
    }

    // byte offsets of the neighbour columns (wave-uniform, scalar registers)
    uint32_t coff[C][KH];
    ACC ed0[C], edU[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        // columns past ndn (last block) repeat column ndn-1 and are not stored
        const int idn = idn0 + c < ndn ? idn0 + c : ndn - 1;
        const Hop *hr = dtab + (long long)idn * op.dn.K;
#pragma unroll
        for (int k = 0; k < KH; ++k) {
            const Hop h = gload(hr + k);
            coff[c][k] = h.col * (uint32_t)nup * (uint32_t)sizeof(T);
            if (threadIdx.x == 0) s_dval[k].v[c] = (ACC)h.val;
        }
        const ACC e = energy_of<ACC>(op.dn, idn);
        ed0[c] = e;
        edU[c] = gload(op.dn.mask + idn) ? e + (ACC)op.U : e;
    }
    // stage the C columns row-interleaved, xs[i*C + c] = x(i, idn0 + c)
    for (int i0 = threadIdx.x; i0 < nup; i0 += kStage * kHvBlock) {
        T v[C][kStage];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int idn = idn0 + c < ndn ? idn0 + c : ndn - 1;
            const T *xc = x + (long long)idn * nup;
#pragma unroll
            for (int r = 0; r < kStage; ++r) {
                const int i = i0 + r * kHvBlock;
                v[c][r] = gload(xc + (i < nup ? i : nup - 1));
            }
        }
#pragma unroll
        for (int r = 0; r < kStage; ++r) {
            const int i = i0 + r * kHvBlock;
            if (i < nup) {
                Row<T, C> w;
#pragma unroll
                for (int c = 0; c < C; ++c) w.v[c] = v[c][r];
                *reinterpret_cast<Row<T, C> *>(xs + i * C) = w;
            }
        }
    }
    __syncthreads();

    double sums[3] = {0.0, 0.0, 0.0};
    double sx = 1.0, bsp = 0.0;
    if constexpr (MODE == HV_LANCZOS) {
        if (!lanczos_scales<kHvBlock>(a, sx, bsp, lead)) return;
    }
    T *y = static_cast<T *>(a.y);
    const T *pv = static_cast<const T *>(a.p);
    const bool has_p = MODE == HV_LANCZOS && pv != nullptr;
    for (int i0 = threadIdx.x; i0 < nup; i0 += U * kHvBlock) {
        int iu[U];
        bool live[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * kHvBlock;
            live[u] = i < nup;
            iu[u] = live[u] ? i : i0;  // a dead row recomputes row i0 and is not stored
        }
        // all loads of the U rows first
        Row<uint32_t, KS> hr[U];
        T g[U][C][KH];
        ACC eu[U];
        bool mu[U];
        T old[U][C];
#pragma unroll
        for (int u = 0; u < U; ++u)
            hr[u] = gload(reinterpret_cast<const Row<uint32_t, KS> *>(op.up.hrec + (long long)iu[u] * KS));
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t ib = (uint32_t)iu[u] * (uint32_t)sizeof(T);
#pragma unroll
            for (int c = 0; c < C; ++c)
#pragma unroll
                for (int k = 0; k < KH; ++k) g[u][c][k] = gload(reinterpret_cast<const T *>(xb + (coff[c][k] + ib)));
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            eu[u] = energy_of<ACC>(op.up, iu[u]);
            mu[u] = gload(op.up.mask + iu[u]) != 0;
        }
        if constexpr (MODE == HV_ACCUM || MODE == HV_LANCZOS) {
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const int idn = idn0 + c < ndn ? idn0 + c : ndn - 1;
                    const long long e = (long long)idn * nup + iu[u];
                    if constexpr (MODE == HV_ACCUM) old[u][c] = gload(y + e);
                    else old[u][c] = has_p ? gload(pv + e) : T(0);
                }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const Row<T, C> xr = *reinterpret_cast<const Row<T, C> *>(xs + iu[u] * C);
            ACC acc[C];
#pragma unroll
            for (int c = 0; c < C; ++c) acc[c] = (eu[u] + (mu[u] ? edU[c] : ed0[c])) * (ACC)xr.v[c];
#pragma unroll
            for (int k = 0; k < KH; ++k) {
                const uint32_t w = hr[u].v[k];
                const ACC val = s_v[w >> 16];
                const Row<T, C> src = *reinterpret_cast<const Row<T, C> *>(xs + (w & 0xFFFFu) * C);
#pragma unroll
                for (int c = 0; c < C; ++c) acc[c] += val * (ACC)src.v[c];
            }
#pragma unroll
            for (int k = 0; k < KH; ++k) {
                const Row<ACC, C> dv = s_dval[k];
#pragma unroll
                for (int c = 0; c < C; ++c) acc[c] += dv.v[c] * (ACC)g[u][c][k];
            }
            if (!live[u]) continue;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                if (idn0 + c >= ndn) break;
                const long long e = (long long)(idn0 + c) * nup + iu[u];
                if constexpr (MODE == HV_PLAIN) {
                    gstore(y + e, (T)acc[c]);
                } else if constexpr (MODE == HV_ACCUM) {
                    gstore(y + e, (T)((ACC)old[u][c] + acc[c]));
                } else if constexpr (MODE == HV_LANCZOS) {
                    double r = sx * (double)acc[c];
                    if (has_p) r -= bsp * (double)old[u][c];
                    const T rt = (T)r;
                    gstore(y + e, rt);
                    sums[0] += (double)xr.v[c] * (double)rt;
                } else {
                    const double xv = (double)xr.v[c], hv = (double)acc[c];
                    sums[0] += xv * hv;
                    sums[1] += hv * hv;
                    sums[2] += xv * xv;
                }
            }
        }
    }
    if constexpr (MODE == HV_LANCZOS) {
        double mine[1] = {sums[0]};
        block_partial_at<1, kHvBlock>(mine, a.partial, slot);  // alpha_j = s_j * sum (next kernel)
    } else if constexpr (MODE == HV_RQ) {
        block_partial_at<3, kHvBlock>(sums, a.partial, slot);
    }
}


}  // namespace edg
