// H.v entry point that allocates nothing (edg_hx_noalloc, include/edgpu.h).
//
// The per-spin basis and hop tables that edg_hx keeps in context-owned
// device memory live here in one file-scope device array, `scratch`, of
// kScratchBytes.  Every GPU the library runs on holds its own copy from
// module load until the process ends.  Host code never uses the array's
// name as an address: at the entry point's first call on a device,
// hipGetSymbolAddress gives that device's copy, which is kept per device
// and passed to the kernels.  All work is queued on the caller's stream
// (kernels only: no allocation, memset or copy), so a caller may capture
// the call in a HIP graph.  Calls overlapping on one GPU share its array
// and race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/edgpu.h"
#include "edgpu_internal.h"

namespace edg {

namespace {

constexpr size_t kScratchBytes = size_t(16) << 20;
constexpr size_t kScratchAlign = 256;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
unsigned char *g_scratch[kMaxDevices] = {};  // device address of each device's copy
int g_ncu[kMaxDevices] = {};
thread_local std::string g_err;

int nfail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

long long binom(int n, int k)
{
    if (k < 0 || k > n) return 0;
    long long r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

// Largest number of hybridization hops of one spin over the configurations
// of ne electrons on norb impurity + nbath bath sites, all V nonzero: a
// configuration with a occupied impurity orbitals has a (nbath - (ne - a))
// + (norb - a)(ne - a) pairs of unlike occupation.
int kmax_bound(int norb, int nbath, int ne)
{
    int best = 0;
    for (int a = 0; a <= norb; ++a) {
        const int b = ne - a;
        if (b < 0 || b > nbath) continue;
        const int h = a * (nbath - b) + (norb - a) * b;
        best = h > best ? h : best;
    }
    return best;
}

// This device's copy of the scratch array (looked up once per device, with
// the device current) and its compute-unit count.
int scratch_of(int dev, unsigned char **p, int *ncu)
{
    if (dev < 0 || dev >= kMaxDevices) return nfail(EDG_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *q = nullptr;
        const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !q) {
            (void)hipGetLastError();
            return nfail(EDG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        }
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
            (void)hipGetLastError();
            return nfail(EDG_ERR_HIP, "hipDeviceGetAttribute failed");
        }
        g_scratch[dev] = static_cast<unsigned char *>(q);
        g_ncu[dev] = n;
    }
    *p = g_scratch[dev];
    *ncu = g_ncu[dev];
    return 0;
}

// Bump layout of the tables inside the scratch array; base == nullptr only
// measures.
struct Carve {
    unsigned char *base;
    size_t off = 0;
    void *take(size_t bytes)
    {
        off = (off + kScratchAlign - 1) / kScratchAlign * kScratchAlign;
        void *p = base ? base + off : nullptr;
        off += bytes < 16 ? 16 : bytes;
        return p;
    }
};

// Layout of one (spin, N) table set, the arrays of edg_hx's context tables
// (pattern, FP32/FP64 hop lists, counts, energies, masks, orbital moves and
// the compressed single-slot records).
void layout(Carve &cv, SpinTab &t, int norb, int nbath, int ns, int ne, bool compress)
{
    t = SpinTab{};
    t.n = (int)binom(ns, ne);
    t.K = norb * nbath > kMinSlots ? norb * nbath : kMinSlots;
    const size_t n = t.n, K = t.K;
    t.pattern = (uint32_t *)cv.take(n * 4);
    t.hf = (HopF *)cv.take(K * n * sizeof(HopF));
    t.hd = (HopD *)cv.take(K * n * sizeof(HopD));
    t.cnt = (uint8_t *)cv.take(n);
    t.maxcnt = (int *)cv.take(sizeof(int));
    t.energy = (double *)cv.take(n * 8);
    t.energyf = (float *)cv.take(n * 4);
    t.mask = (uint8_t *)cv.take(n);
    if (norb >= 2) t.orb = (uint32_t *)cv.take((size_t)norb * norb * n * 4);
    if (compress) {
        t.ccol = (uint16_t *)cv.take(n * 16 * 2);
        t.cidx = (uint8_t *)cv.take(n * 16);
        t.hrec = (uint32_t *)cv.take(n * 16 * 4);
    }
}

}  // namespace

const char *noalloc_last_error() { return g_err.c_str(); }

}  // namespace edg

using namespace edg;

extern "C" int edg_hx_noalloc(const edg_params *p, const double *em, const double *ek, const double *vmk,
                              int ne_up, int ne_dn, const void *x, void *y, int py, void *stream)
{
    if (!p || !em || !x || !y) return nfail(EDG_ERR_ARG, "edg_hx_noalloc: p, em, x and y must not be NULL");
    if (p->norb < 1 || p->nbath < 0 || p->norb + p->nbath > 15 || p->norb > kMaxNorb ||
        p->norb * p->nbath > kMaxHops || (p->precision != EDG_FP32 && p->precision != EDG_FP64))
        return nfail(EDG_ERR_ARG, "edg_hx_noalloc: invalid norb, nbath or precision");
    if (p->nbath > 0 && (!ek || !vmk)) return nfail(EDG_ERR_ARG, "edg_hx_noalloc: ek and vmk must not be NULL");
    const int norb = p->norb, nbath = p->nbath, ns = norb + nbath;
    if (ne_up < 0 || ne_dn < 0 || ne_up > ns || ne_dn > ns)
        return nfail(EDG_ERR_ARG, "edg_hx_noalloc: invalid sector (" + std::to_string(ne_up) + "," +
                                      std::to_string(ne_dn) + ")");
    const int km[2] = {kmax_bound(norb, nbath, ne_up), kmax_bound(norb, nbath, ne_dn)};
    const int ne[2] = {ne_up, ne_dn};
    bool comp[2];
    for (int s = 0; s < 2; ++s) comp[s] = norb <= 2 && km[s] <= 16 && binom(ns, ne[s]) <= 65536;
    // measure, check the limit, then place
    Carve probe{nullptr};
    SpinTab tab[2];
    for (int s = 0; s < 2; ++s) layout(probe, tab[s], norb, nbath, ns, ne[s], comp[s]);
    if (probe.off > kScratchBytes)
        return nfail(EDG_ERR_ARG, "edg_hx_noalloc: the tables of sector (" + std::to_string(ne_up) + "," +
                                      std::to_string(ne_dn) + ") need " + std::to_string(probe.off) +
                                      " bytes; the no-allocation scratch holds " + std::to_string(kScratchBytes));
    const size_t es = p->precision == EDG_FP32 ? 4 : 8;
    const long long D = binom(ns, ne_up) * binom(ns, ne_dn);
    const char *xb = static_cast<const char *>(x);
    char *yb = static_cast<char *>(y);
    if (xb < yb + D * es && yb < xb + D * es) return nfail(EDG_ERR_ARG, "edg_hx_noalloc: x and y overlap");
    const double Up = p->Up < 0 ? p->U - 2.0 * p->J : p->Up;  // dmft_params.f90:90-92
    const double Jp = p->Jp < 0 ? p->J : p->Jp;                // dmft_params.f90:94-96
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return nfail(EDG_ERR_HIP, "edg_hx_noalloc: no current HIP device");
    }
    unsigned char *base = nullptr;
    int ncu = 0;
    if (int rc = scratch_of(dev, &base, &ncu)) return rc;
    Carve cv{base};
    hipStream_t st = static_cast<hipStream_t>(stream);
    const double UpmJ = Up - p->J;
    for (int s = 0; s < 2; ++s) {
        layout(cv, tab[s], norb, nbath, ns, ne[s], comp[s]);
        const double *em_s = em + s * norb;
        const double *ek_s = nbath ? ek + s * nbath : nullptr;
        const double *v_s = nbath ? vmk + (size_t)s * nbath * norb : nullptr;
        if (launch_spin_table(st, ns, norb, nbath, ne[s], em_s, ek_s, v_s, p->mu, UpmJ, tab[s]))
            return nfail(EDG_ERR_HIP, "edg_hx_noalloc: spin table launch failed");
        tab[s].kmax = km[s];
        tab[s].vsrc = v_s;
        if (comp[s]) {
            tab[s].KC = km[s] <= 8 ? 8 : 16;
            if (launch_compress(st, tab[s])) return nfail(EDG_ERR_HIP, "edg_hx_noalloc: compress launch failed");
        }
    }
    HvOp op{};
    op.nup = tab[0].n;
    op.ndn = tab[1].n;
    op.norb = norb;
    op.nvup = norb * nbath;
    op.ncu = ncu;
    op.sched = 0;
    op.up = tab[0];
    op.dn = tab[1];
    op.U = p->U;
    op.Up = Up;
    op.Jp = Jp;
    HvArgs a{};
    a.x = x;
    a.y = y;
    a.accumulate = py;
    // the default schedule of a context (edg_create)
    int cols;
    bool acc64;
    default_schedule(norb, p->precision == EDG_FP32, &cols, &acc64);
    const int rc = es == 4 ? launch_hv<float>(st, op, HV_PLAIN, cols, acc64, a)
                           : launch_hv<double>(st, op, HV_PLAIN, cols, acc64, a);
    if (rc < 0) return nfail(EDG_ERR_HIP, "edg_hx_noalloc: H.v launch failed");
    return EDG_OK;
}
