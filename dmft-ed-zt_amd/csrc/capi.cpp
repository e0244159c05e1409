// C-ABI layer of the MI355X ED impurity solver: context, device memory,
// per-spin table cache and the Lanczos / ground-state / Green's-function
// drivers.  Entry points are declared in include/edgpu.h.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/edgpu.h"
#include "devmem.h"
#include "edgpu_internal.h"

namespace edg {
int tridiag_lowest(int n, const double *d, const double *e, double *lambda, double *z);
const char *noalloc_last_error();
}

using namespace edg;

namespace {

constexpr long long kPad = 64;  // element alignment of stored vectors
constexpr long long kPartialHalf = 65536;  // doubles per half of the partial-sum buffer

long long round_up(long long n, long long m) { return (n + m - 1) / m * m; }

long long binom(int n, int k)
{
    if (k < 0 || k > n) return 0;
    long long r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct TabEntry {
    SpinTab t;
    std::vector<void *> allocs;
    unsigned long long version = 0;
};

// Device scalars of one Lanczos run.
struct Scalars {
    double *alpha, *beta, *scale, *misc;  // misc[0] = |v_init|, misc[1..3] RQ sums
    int *flags;                            // flags[0] = stop, flags[1] = nstep
    FusedCoef *fcoef;                      // [maxn + 1] coefficients of the fused steps
};

}  // namespace

struct edg_ctx {
    edg_params P{};
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    unsigned long long budget = 0;
    int cols = 4;
    bool acc64 = false;
    bool fused = false;  // single-orbital Lanczos steps without the axpy pass (HV_FUSED)
    int reorth = 0;      // reorthogonalization of the Green's-function chains: 0 none, 1 partial, 2 full
    long long ro_events = 0, ro_steps = 0;
    int ncu = 0;

    std::vector<double> em, ek, vmk;
    unsigned long long version = 0;
    std::map<std::pair<int, int>, TabEntry> tabs;
    DevBuf dparams;  // em | ek | vmk on device

    double *partial = nullptr;
    Scalars sc{};
    void *scal_mem = nullptr;

    DevBuf basis, ring, gs[2], stage, tmp, coefbuf, gbuf, omegabuf, mapbuf, sigbuf;
    int gs_slot = 0;

    bool has_gs = false;
    int gs_nu = 0, gs_nd = 0, gs_isector = 0, gs_nstep = 0;
    long long gs_dim = 0;
    double e0 = 0, e0_ritz = 0, residual = 0;

    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    long long n_hv = 0, n_steps = 0;
    bool have_gs_ev = false, have_gf_ev = false;

    // A second stream with its own partial sums, scalars and ring: the c+
    // and c chains of one (orbital, spin) run concurrently, the second
    // filling the first one's kernel tails and host round trips.
    hipStream_t s2 = nullptr;
    double *partial2 = nullptr;
    Scalars sc2{};
    void *scal_mem2 = nullptr;
    DevBuf ring2, mapbuf2;
    DevBuf robuf, robuf2;  // reorthogonalization state of the two lanes
    hipEvent_t ev2[2] = {nullptr, nullptr};  // [0] main stream ready, [1] second stream done
};

namespace {

int fail(edg_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

// The stream and per-run device buffers of one Lanczos run in flight.
struct Lane {
    hipStream_t stream;
    double *partial;
    const Scalars *sc;
    DevBuf *ring, *mapbuf, *robuf;
};

Lane lane_of(edg_ctx *c, int k)
{
    if (k == 0) return Lane{c->stream, c->partial, &c->sc, &c->ring, &c->mapbuf, &c->robuf};
    return Lane{c->s2, c->partial2, &c->sc2, &c->ring2, &c->mapbuf2, &c->robuf2};
}

#define HIPCHK(c, call)                                                                 \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(c, EDG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// hipMalloc that, when the device is full, first returns the idle memory
// chunks closed batched queues left on it and tries again.
bool dev_malloc(void **p, size_t bytes, int device)
{
    *p = nullptr;
    if (hipMalloc(p, bytes) == hipSuccess) return true;
    (void)hipGetLastError();
    *p = nullptr;
    (void)devmem_release_cache(device);
    if (hipMalloc(p, bytes) == hipSuccess) return true;
    (void)hipGetLastError();
    *p = nullptr;
    return false;
}

int ensure(edg_ctx *c, DevBuf &b, size_t bytes)
{
    if (b.bytes >= bytes) return 0;
    if (b.p) {  // either stream may still use it
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (c->s2) HIPCHK(c, hipStreamSynchronize(c->s2));
        HIPCHK(c, hipFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
    }
    if (!dev_malloc(&b.p, bytes, c->device)) {
        return fail(c, EDG_ERR_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    }
    b.bytes = bytes;
    return 0;
}

size_t elem_size(const edg_ctx *c) { return c->P.precision == EDG_FP32 ? 4 : 8; }

int check_sector(edg_ctx *c, int nu, int nd)
{
    const int ns = c->P.norb + c->P.nbath;
    if (nu < 0 || nd < 0 || nu > ns || nd > ns)
        return fail(c, EDG_ERR_ARG, "invalid sector (" + std::to_string(nu) + "," + std::to_string(nd) +
                                        ") for nsite=" + std::to_string(ns));
    return 0;
}

// Install em/ek/vmk; bump the version (invalidating tables) when they change.
int set_params(edg_ctx *c, const double *em, const double *ek, const double *vmk)
{
    const int norb = c->P.norb, nbath = c->P.nbath;
    if (!em || (!ek && nbath) || (!vmk && nbath)) return fail(c, EDG_ERR_ARG, "em/ek/vmk must not be NULL");
    std::vector<double> a(em, em + 2 * norb), b, v;
    if (nbath) {
        b.assign(ek, ek + 2 * nbath);
        v.assign(vmk, vmk + 2 * norb * nbath);
    }
    if (c->version && a == c->em && b == c->ek && v == c->vmk) return 0;
    c->em = a;
    c->ek = b;
    c->vmk = v;
    ++c->version;
    const size_t n = a.size() + b.size() + v.size();
    if (int rc = ensure(c, c->dparams, n * sizeof(double))) return rc;
    std::vector<double> all;
    all.insert(all.end(), a.begin(), a.end());
    all.insert(all.end(), b.begin(), b.end());
    all.insert(all.end(), v.begin(), v.end());
    HIPCHK(c, hipMemcpyAsync(c->dparams.p, all.data(), n * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int get_tab(edg_ctx *c, int spin, int ne, const SpinTab **out)
{
    auto key = std::make_pair(spin, ne);
    TabEntry &te = c->tabs[key];
    const int norb = c->P.norb, nbath = c->P.nbath, ns = norb + nbath;
    if (te.allocs.empty()) {
        SpinTab &t = te.t;
        t.n = (int)binom(ns, ne);
        t.K = norb * nbath > kMinSlots ? norb * nbath : kMinSlots;
        auto al = [&](void **p, size_t bytes) -> int {
            if (!dev_malloc(p, bytes < 16 ? 16 : bytes, c->device)) return fail(c, EDG_ERR_NOMEM, "table allocation failed");
            te.allocs.push_back(*p);
            return 0;
        };
        const size_t n = t.n, K = t.K;
        int rc = 0;
        rc |= al((void **)&t.pattern, n * 4);
        rc |= al((void **)&t.hf, K * n * sizeof(HopF));
        rc |= al((void **)&t.hd, K * n * sizeof(HopD));
        rc |= al((void **)&t.cnt, n);
        rc |= al((void **)&t.maxcnt, sizeof(int));
        rc |= al((void **)&t.energy, n * 8);
        rc |= al((void **)&t.energyf, n * 4);
        rc |= al((void **)&t.mask, n);
        if (norb >= 2) rc |= al((void **)&t.orb, (size_t)norb * norb * n * 4);
        if (rc) return rc;
    }
    if (te.version != c->version) {
        const double *dp = (const double *)c->dparams.p;
        const double *em_s = dp + (spin - 1) * norb;
        const double *ek_s = dp + 2 * norb + (spin - 1) * nbath;
        const double *v_s = dp + 2 * norb + 2 * nbath + (size_t)(spin - 1) * nbath * norb;
        const double UpmJ = c->P.Up - c->P.J;
        HIPCHK(c, hipMemsetAsync(te.t.maxcnt, 0, sizeof(int), c->stream));
        if (launch_spin_table(c->stream, ns, norb, nbath, ne, em_s, ek_s, v_s, c->P.mu, UpmJ, te.t))
            return fail(c, EDG_ERR_HIP, "spin table kernel launch failed");
        int km = 0;
        HIPCHK(c, hipMemcpyAsync(&km, te.t.maxcnt, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        te.t.kmax = km;
        te.t.vsrc = v_s;
        if (norb <= 2 && km <= 16 && te.t.n <= 65536) {  // compressed tables (16-bit ranks)
            te.t.KC = km <= 8 ? 8 : 16;
            if (!te.t.ccol) {
                auto al2 = [&](void **p, size_t bytes) -> int {
                    if (!dev_malloc(p, bytes, c->device)) return fail(c, EDG_ERR_NOMEM, "table allocation failed");
                    te.allocs.push_back(*p);
                    return 0;
                };
                if (int rc = al2((void **)&te.t.ccol, (size_t)te.t.n * 16 * 2)) return rc;
                if (int rc = al2((void **)&te.t.cidx, (size_t)te.t.n * 16)) return rc;
                if (int rc = al2((void **)&te.t.hrec, (size_t)te.t.n * 16 * 4)) return rc;
            }
            if (launch_compress(c->stream, te.t)) return fail(c, EDG_ERR_HIP, "compress launch failed");
        } else {
            te.t.KC = 0;
        }
        te.version = c->version;
    }
    *out = &te.t;
    return 0;
}

int make_op(edg_ctx *c, int nu, int nd, HvOp *op)
{
    if (int rc = check_sector(c, nu, nd)) return rc;
    const SpinTab *tu, *td;
    if (int rc = get_tab(c, 1, nu, &tu)) return rc;
    if (int rc = get_tab(c, 2, nd, &td)) return rc;
    op->nup = tu->n;
    op->ndn = td->n;
    op->norb = c->P.norb;
    op->nvup = c->P.norb * c->P.nbath;
    op->ncu = c->ncu;
    op->sched = 0;
    op->up = *tu;
    op->dn = *td;
    op->U = c->P.U;
    op->Up = c->P.Up;
    op->Jp = c->P.Jp;
    return 0;
}

HvArgs base_args(edg_ctx *c)
{
    HvArgs a{};
    a.partial = c->partial;
    return a;
}

// Run up to maxn Lanczos steps (lanczos_solver.f90:92-173) on vectors held
// either as a stored basis (slot j at base + j*stride, maxn+1 slots) or a
// ring of 3 slots.  Slot 0 holds the unnormalized start vector.  The steps
// are queued on the lane's stream without waiting; lanczos_finish waits for
// them and fills the host arrays a, b (maxn entries, zero-padded) and the
// scales s_j (if not null); *vinit_norm receives |v_init|.
// Layout of a lane's reorthogonalization buffer (doubles): omega 3 rows,
// coef, flags (8 ints), then the partial dots (kReorthDots per vector).
int reorth_ld(int maxn) { return (maxn + 2 + 7) / 8 * 8; }
size_t reorth_bytes(int maxn) { return sizeof(double) * (size_t)reorth_ld(maxn) * (4 + kReorthDots) + 64; }

template <typename T>
int lanczos_enqueue(edg_ctx *c, const Lane &L, const HvOp &op, T *base, long long stride, bool stored,
                    int maxn, T *rbase, bool ro = false)
{
    const long long D = (long long)op.nup * op.ndn;
    const Scalars &S = *L.sc;
    auto slot = [&](int j) -> T * { return base + (stored ? (long long)j : (long long)(j % 3)) * stride; };
    HIPCHK(c, hipMemsetAsync(S.alpha, 0, sizeof(double) * maxn, L.stream));
    HIPCHK(c, hipMemsetAsync(S.beta, 0, sizeof(double) * (maxn + 1), L.stream));
    HIPCHK(c, hipMemsetAsync(S.scale, 0, sizeof(double) * (maxn + 1), L.stream));
    int fl[2] = {0, maxn};
    HIPCHK(c, hipMemcpyAsync(S.flags, fl, sizeof(fl), hipMemcpyHostToDevice, L.stream));
    if (launch_norm<T>(L.stream, slot(0), D, L.partial, &S.misc[0], &S.scale[0], nullptr))
        return fail(c, EDG_ERR_HIP, "norm launch failed");
    if (rbase && !ro && c->fused && hv_fused_supported(op, c->cols, sizeof(T))) {
        // one H.v launch per step on r'_{j-1} (ring of two at rbase) and its
        // finisher; u_j goes to slot j (stored basis) or j % 3 (ring)
        auto rslot = [&](int j) -> T * { return rbase + (long long)(j & 1) * stride; };
        if (launch_fused_start(L.stream, S.scale, S.fcoef)) return fail(c, EDG_ERR_HIP, "fused start launch failed");
        for (int j = 0; j < maxn; ++j) {
            HvArgs ha{};
            ha.x = j == 0 ? slot(0) : rslot(j - 1);
            ha.y = rslot(j);
            ha.p = j >= 1 ? slot(j - 1) : nullptr;
            ha.p2 = j >= 2 ? slot(j - 2) : nullptr;
            ha.u = j >= 1 ? slot(j) : nullptr;
            ha.partial = L.partial;
            ha.step = j;
            ha.alpha = S.alpha;
            ha.beta = S.beta;
            ha.scale = S.scale;
            ha.stop = S.flags;
            ha.nstep = S.flags + 1;
            ha.fcoef = S.fcoef;
            const int hv_n = launch_hv<T>(L.stream, op, HV_FUSED, c->cols, c->acc64, ha);
            if (hv_n < 0) return fail(c, EDG_ERR_HIP, "fused H.v launch failed");
            ++c->n_hv;
            ++c->n_steps;
            if (launch_fused_finish(L.stream, L.partial, hv_n, j, S.alpha, S.beta, S.scale, S.flags,
                                    S.flags + 1, S.fcoef))
                return fail(c, EDG_ERR_HIP, "fused finisher launch failed");
        }
        return 0;
    }
    // H.v partials and axpy partials live in separate halves of the buffer:
    // step j's H.v reads step j-1's axpy partials while writing its own
    double *hv_part = L.partial, *ax_part = L.partial + kPartialHalf;
    int ax_n = 0;
    ReorthArgs ra{};
    if (ro) {  // a stored basis, reorthogonalized after each axpy (reorth.hip)
        if (!stored) return fail(c, EDG_ERR_ARG, "reorthogonalization needs the stored Krylov basis");
        if (int rc = ensure(c, *L.robuf, reorth_bytes(maxn))) return rc;
        const int ld = reorth_ld(maxn);
        double *rb = (double *)L.robuf->p;
        ra.mode = c->reorth;
        ra.ld = ld;
        ra.eps = sizeof(T) == 4 ? 5.9604644775390625e-08 : 1.1102230246251565e-16;  // 2^-24, 2^-53
        ra.alpha = S.alpha;
        ra.beta = S.beta;
        ra.scale = S.scale;
        ra.ax_part = ax_part;
        ra.omega = rb;
        ra.coef = rb + 3 * (size_t)ld;
        ra.flags = (int *)(rb + 4 * (size_t)ld);
        ra.dpart = rb + 4 * (size_t)ld + 4;
        ra.stop = S.flags;
        static const double kOne = 1.0;  // omega_{0,0}
        HIPCHK(c, hipMemsetAsync(rb, 0, sizeof(double) * (4 * (size_t)ld + 4), L.stream));
        HIPCHK(c, hipMemcpyAsync(rb, &kOne, sizeof(double), hipMemcpyHostToDevice, L.stream));
    }
    for (int j = 0; j < maxn; ++j) {
        HvArgs ha{};
        ha.x = slot(j);
        ha.y = slot(j + 1);
        ha.p = j > 0 ? slot(j - 1) : nullptr;
        ha.partial = hv_part;
        ha.step = j;
        ha.prev_part = ax_part;
        ha.prev_n = ax_n;
        ha.scale = S.scale;
        ha.beta = S.beta;
        ha.stop = S.flags;
        ha.nstep = S.flags + 1;
        const int hv_n = launch_hv<T>(L.stream, op, HV_LANCZOS, c->cols, c->acc64, ha);
        if (hv_n < 0) return fail(c, EDG_ERR_HIP, "H.v launch failed");
        ++c->n_hv;
        ++c->n_steps;
        if (j == maxn - 1) {
            if (launch_alpha(L.stream, hv_part, hv_n, &S.scale[j], &S.alpha[j], S.flags))
                return fail(c, EDG_ERR_HIP, "alpha launch failed");
            break;
        }
        ax_n = launch_axpy_norm<T>(L.stream, slot(j + 1), slot(j), D, hv_part, hv_n, S.scale, S.alpha,
                                   j, ax_part, S.flags);
        if (ax_n < 0) return fail(c, EDG_ERR_HIP, "axpy launch failed");
        if (ro) {
            ra.step = j;
            ra.ax_n = ax_n;
            if (launch_reorth<T>(L.stream, ra, slot(j + 1), base, stride, D, kReorthDots))
                return fail(c, EDG_ERR_HIP, "reorthogonalization launch failed");
        }
    }
    return 0;
}

int lanczos_finish(edg_ctx *c, const Lane &L, int maxn, int *nstep, double *a, double *b, double *scales,
                   double *vinit_norm)
{
    const Scalars &S = *L.sc;
    std::vector<double> ha(maxn), hb(maxn + 1), hs(maxn + 1);
    int hf[2];
    double nv;
    HIPCHK(c, hipMemcpyAsync(ha.data(), S.alpha, sizeof(double) * maxn, hipMemcpyDeviceToHost, L.stream));
    HIPCHK(c, hipMemcpyAsync(hb.data(), S.beta, sizeof(double) * (maxn + 1), hipMemcpyDeviceToHost, L.stream));
    HIPCHK(c, hipMemcpyAsync(hs.data(), S.scale, sizeof(double) * (maxn + 1), hipMemcpyDeviceToHost, L.stream));
    HIPCHK(c, hipMemcpyAsync(hf, S.flags, sizeof(hf), hipMemcpyDeviceToHost, L.stream));
    HIPCHK(c, hipMemcpyAsync(&nv, &S.misc[0], sizeof(double), hipMemcpyDeviceToHost, L.stream));
    HIPCHK(c, hipStreamSynchronize(L.stream));
    const int ns = hf[1];
    *nstep = ns;
    for (int k = 0; k < maxn; ++k) {
        a[k] = k < ns ? ha[k] : 0.0;
        b[k] = (k >= 1 && k <= ns) ? hb[k] : 0.0;  // b(nstep+1) is kept when the chain broke down
    }
    if (ns < maxn && ns >= 1) b[ns] = hb[ns];
    b[0] = 0.0;
    if (scales)
        for (int k = 0; k <= maxn; ++k) scales[k] = hs[k];
    if (vinit_norm) *vinit_norm = nv;
    return 0;
}

template <typename T>
int run_lanczos(edg_ctx *c, const Lane &L, const HvOp &op, T *base, long long stride, bool stored, int maxn,
                T *rbase, int *nstep, double *a, double *b, double *scales, double *vinit_norm)
{
    if (int rc = lanczos_enqueue<T>(c, L, op, base, stride, stored, maxn, rbase)) return rc;
    return lanczos_finish(c, L, maxn, nstep, a, b, scales, vinit_norm);
}

template <typename T>
int upload(edg_ctx *c, T *dst, const double *src, long long n)
{
    if (int rc = ensure(c, c->stage, n * sizeof(double))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->stage.p, src, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (launch_from_f64<T>(c->stream, dst, (const double *)c->stage.p, n))
        return fail(c, EDG_ERR_HIP, "convert launch failed");
    return 0;
}

template <typename T>
int download(edg_ctx *c, double *dst, const T *src, long long n)
{
    if (int rc = ensure(c, c->stage, n * sizeof(double))) return rc;
    if (launch_to_f64<T>(c->stream, (double *)c->stage.p, src, n))
        return fail(c, EDG_ERR_HIP, "convert launch failed");
    HIPCHK(c, hipMemcpyAsync(dst, c->stage.p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int admit(edg_ctx *c, size_t bytes)
{
    if (c->budget && bytes > c->budget)
        return fail(c, EDG_ERR_ADMISSION, "request needs " + std::to_string(bytes) +
                                              " device bytes, budget is " + std::to_string(c->budget));
    return 0;
}

// diag_lanczos (lanczos_solver.f90:20-90) on one sector with the stored
// Krylov basis: Lanczos, lowest Ritz pair, eigenvector = sum_j z_j q_j,
// then E0 as the FP64 Rayleigh quotient of the stored vector.
template <typename T>
int diag_sector(edg_ctx *c, int nu, int nd, unsigned long long seed, T *gs_out, double *e0_ritz,
                double *e0, double *residual, int *nstep_out)
{
    HvOp op;
    if (int rc = make_op(c, nu, nd, &op)) return rc;
    const int maxn = c->P.maxnstep;
    const long long D = (long long)op.nup * op.ndn, Dp = round_up(D, kPad);
    // the stored Krylov basis (maxn + 1 slots), and the two residual slots
    // of the fused step when it is on
    const size_t need = (size_t)(maxn + (c->fused ? 3 : 1)) * Dp * sizeof(T);
    if (int rc = admit(c, need + 2 * Dp * sizeof(T))) return rc;
    if (int rc = ensure(c, c->basis, need)) return rc;
    T *base = (T *)c->basis.p;
    if (launch_start<T>(c->stream, base, D, seed)) return fail(c, EDG_ERR_HIP, "start launch failed");
    int nstep;
    std::vector<double> a(maxn), b(maxn), s(maxn + 1);
    if (int rc = run_lanczos<T>(c, lane_of(c, 0), op, base, Dp, true, maxn,
                                c->fused ? base + (long long)(maxn + 1) * Dp : nullptr, &nstep, a.data(),
                                b.data(), s.data(), nullptr))
        return rc;
    std::vector<double> z(nstep);
    double lam;
    if (tridiag_lowest(nstep, a.data(), b.data() + 1, &lam, z.data()))
        return fail(c, EDG_ERR_SOLVER, "tridiagonal eigensolve failed");
    std::vector<double> coef(nstep);
    for (int j = 0; j < nstep; ++j) coef[j] = z[j] * s[j];
    if (int rc = ensure(c, c->coefbuf, sizeof(double) * 4 * (maxn + 8))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->coefbuf.p, coef.data(), sizeof(double) * nstep, hipMemcpyHostToDevice,
                             c->stream));
    double *misc = c->sc.misc;
    if (launch_combine<T>(c->stream, gs_out, base, Dp, nstep, (const double *)c->coefbuf.p, D,
                          c->partial, &misc[4], &misc[5]))
        return fail(c, EDG_ERR_HIP, "combine launch failed");
    if (launch_scale<T>(c->stream, gs_out, D, &misc[5])) return fail(c, EDG_ERR_HIP, "scale launch failed");
    HvArgs ha = base_args(c);
    ha.x = gs_out;
    ha.result = &misc[1];
    if (launch_hv<T>(c->stream, op, HV_RQ, c->cols, true, ha) < 0) return fail(c, EDG_ERR_HIP, "RQ launch failed");
    ++c->n_hv;
    double sums[3];
    HIPCHK(c, hipMemcpyAsync(sums, &misc[1], sizeof(sums), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double xx = sums[2], xHx = sums[0], HH = sums[1];
    const double rq = xHx / xx;
    *e0_ritz = lam;
    *e0 = rq;
    *residual = (HH - 2.0 * rq * xHx + rq * rq * xx) / xx;
    if (*residual < 0) *residual = 0.0;
    *nstep_out = nstep;
    return 0;
}

template <typename T>
int ground_state_t(edg_ctx *c, int nsector, const int *sectors, int *isector, double *e0)
{
    c->has_gs = false;
    const size_t es = sizeof(T);
    double best = HUGE_VAL;
    int bi = 0;
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    for (int i = 0; i < nsector; ++i) {
        const int nu = sectors[2 * i], nd = sectors[2 * i + 1];
        if (int rc = check_sector(c, nu, nd)) return rc;
        const long long D = binom(c->P.norb + c->P.nbath, nu) * binom(c->P.norb + c->P.nbath, nd);
        const int cand = 1 - c->gs_slot;
        if (int rc = ensure(c, c->gs[cand], round_up(D, kPad) * es)) return rc;
        double er, e, res;
        int ns;
        if (int rc = diag_sector<T>(c, nu, nd, c->P.seed + i, (T *)c->gs[cand].p, &er, &e, &res, &ns))
            return rc;
        if (e < best) {  // strict minimum: first lowest wins (ground_state.f90:62)
            best = e;
            bi = i + 1;
            c->gs_slot = cand;
            c->gs_nu = nu;
            c->gs_nd = nd;
            c->gs_dim = D;
            c->e0 = e;
            c->e0_ritz = er;
            c->residual = res;
            c->gs_nstep = ns;
        }
    }
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    c->have_gs_ev = true;
    if (bi == 0) {  // no finite E0 in any sector (NaN comparisons never win)
        *isector = 0;
        *e0 = best;
        return fail(c, EDG_ERR_SOLVER, "no sector gave a finite ground-state energy");
    }
    c->has_gs = true;
    c->gs_isector = bi;
    *isector = bi;
    *e0 = best;
    return 0;
}

// One green_diag chain (dmft_green.f90:202-257, one of G+ / G-): v_init =
// c(+)|gs> in the neighbouring sector, Lanczos, and the coefficient arrays
// with b(1) = |v_init| (:219,:249).  chain_enqueue queues it on a lane;
// chain_finish waits for it and fills the coefficients.
struct ChainRun {
    bool active = false;  // false: the sector does not exist (no weight)
    int maxn = 0;
};

// The sector of a chain; false when it does not exist.
bool chain_sector(const edg_ctx *c, int ispin, int pm, int *nu, int *nd)
{
    const int ns_site = c->P.norb + c->P.nbath;
    *nu = c->gs_nu + (ispin == 1 ? (pm == 1 ? 1 : -1) : 0);
    *nd = c->gs_nd + (ispin == 2 ? (pm == 1 ? 1 : -1) : 0);
    return *nu >= 0 && *nd >= 0 && *nu <= ns_site && *nd <= ns_site;
}

template <typename T>
int chain_enqueue(edg_ctx *c, const Lane &L, int iorb, int ispin, int pm, ChainRun *r)
{
    const int ns_site = c->P.norb + c->P.nbath, maxn = c->P.maxnstep_green;
    r->active = false;
    r->maxn = maxn;
    int nu, nd;
    if (!chain_sector(c, ispin, pm, &nu, &nd)) return 0;
    HvOp op, op0;
    if (int rc = make_op(c, nu, nd, &op)) return rc;
    if (int rc = make_op(c, c->gs_nu, c->gs_nd, &op0)) return rc;
    const long long D = (long long)op.nup * op.ndn, Dp = round_up(D, kPad);
    // ring of three Krylov vectors, and the two residual slots of the fused
    // step when it is on; the whole Krylov basis when it is reorthogonalized
    const bool ro = c->reorth != 0;
    const long long nring = ro ? maxn + 1 : (c->fused ? 5 : 3);
    if (int rc = admit(c, nring * Dp * sizeof(T))) return rc;
    if (int rc = ensure(c, *L.ring, nring * Dp * sizeof(T))) return rc;
    if (int rc = ensure(c, *L.mapbuf, sizeof(uint32_t) * 8192)) return rc;
    T *base = (T *)L.ring->p;
    const SpinTab &tout = ispin == 1 ? op.up : op.dn;
    const int ne_in = ispin == 1 ? c->gs_nu : c->gs_nd;
    if (launch_cmap(L.stream, tout.pattern, tout.n, ns_site, ne_in, iorb - 1, pm, (uint32_t *)L.mapbuf->p))
        return fail(c, EDG_ERR_HIP, "cmap launch failed");
    const int godd = ispin == 2 ? (c->gs_nu & 1) : 0;
    if (launch_apply_c<T>(L.stream, (const T *)c->gs[c->gs_slot].p, base, op.nup, op.ndn, op0.nup, ispin,
                          (const uint32_t *)L.mapbuf->p, godd))
        return fail(c, EDG_ERR_HIP, "apply_c launch failed");
    if (int rc = lanczos_enqueue<T>(c, L, op, base, Dp, ro, maxn, !ro && c->fused ? base + 3 * Dp : nullptr, ro))
        return rc;
    r->active = true;
    return 0;
}

int chain_finish(edg_ctx *c, const Lane &L, const ChainRun &r, double *aa, double *bb, int *nstep)
{
    for (int k = 0; k < r.maxn; ++k) aa[k] = bb[k] = 0.0;
    *nstep = 0;
    if (!r.active) return 0;
    double nv;
    if (int rc = lanczos_finish(c, L, r.maxn, nstep, aa, bb, nullptr, &nv)) return rc;
    bb[0] = nv;
    if (c->reorth && L.robuf->p) {  // reorthogonalized steps of this chain
        int fl[8];
        HIPCHK(c, hipMemcpy(fl, (const char *)L.robuf->p + sizeof(double) * 4 * reorth_ld(r.maxn), sizeof(fl),
                            hipMemcpyDeviceToHost));
        c->ro_events += fl[2];
        c->ro_steps += fl[5];
    }
    return 0;
}

template <typename T>
int cluster_green_t(edg_ctx *c, int nw, const double *omega, double *G_cl, double *coeffs)
{
    if (!c->has_gs) return fail(c, EDG_ERR_STATE, "no ground state held: call edg_ground_state first");
    c->ro_events = c->ro_steps = 0;
    const int norb = c->P.norb, nspin = c->P.nspin, maxn = c->P.maxnstep_green;
    const bool even = !(nspin == 2 || c->gs_nu != c->gs_nd);
    if (int rc = ensure(c, c->omegabuf, sizeof(double) * nw)) return rc;
    if (int rc = ensure(c, c->gbuf, sizeof(double) * 2 * nw * 2)) return rc;
    if (int rc = ensure(c, c->coefbuf, sizeof(double) * 4 * (maxn + 8))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->omegabuf.p, omega, sizeof(double) * nw, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
    // the tables of the ground-state and every chain sector first: building
    // one waits for the main stream, which would serialize the two streams'
    // chains
    {
        HvOp op;
        if (int rc = make_op(c, c->gs_nu, c->gs_nd, &op)) return rc;
    }
    for (int ispin = 1; ispin <= 2; ++ispin)
        for (int pm = 1; pm <= 2; ++pm) {
            int nu, nd;
            HvOp op;
            if (chain_sector(c, ispin, pm, &nu, &nd))
                if (int rc = make_op(c, nu, nd, &op)) return rc;
        }
    std::vector<double> g(2 * nw * 2);
    std::vector<double> cf(4 * maxn);
    for (int iorb = 1; iorb <= norb; ++iorb) {
        double *co[2][4];
        std::vector<double> keep[2];
        for (int ispin = 1; ispin <= 2; ++ispin) {
            const bool run = ispin == 1 || !even || c->P.both_spins;
            keep[ispin - 1].assign(4 * maxn, 0.0);
            for (int q = 0; q < 4; ++q) co[ispin - 1][q] = keep[ispin - 1].data() + q * maxn;
            if (!run) {
                keep[1] = keep[0];  // the reference writes spin-1 coefficients again (:102-109)
                continue;
            }
            // c+ chain on the main stream, c chain on the second one, after
            // everything queued so far (ground state, tables)
            const Lane L0 = lane_of(c, 0), L1 = lane_of(c, 1);
            HIPCHK(c, hipEventRecord(c->ev2[0], c->stream));
            HIPCHK(c, hipStreamWaitEvent(c->s2, c->ev2[0], 0));
            ChainRun rp, rm;
            if (int rc = chain_enqueue<T>(c, L0, iorb, ispin, 1, &rp)) return rc;
            if (int rc = chain_enqueue<T>(c, L1, iorb, ispin, 2, &rm)) return rc;
            int np, nm;
            if (int rc = chain_finish(c, L0, rp, co[ispin - 1][0], co[ispin - 1][1], &np)) return rc;
            if (int rc = chain_finish(c, L1, rm, co[ispin - 1][2], co[ispin - 1][3], &nm)) return rc;
            HIPCHK(c, hipEventRecord(c->ev2[1], c->s2));
            HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev2[1], 0));
            HIPCHK(c, hipMemcpyAsync(c->coefbuf.p, keep[ispin - 1].data(), sizeof(double) * 4 * maxn,
                                     hipMemcpyHostToDevice, c->stream));
            const double *dc = (const double *)c->coefbuf.p;
            double *dg = (double *)c->gbuf.p + (size_t)(ispin - 1) * 2 * nw;
            if (launch_cf(c->stream, nw, (const double *)c->omegabuf.p, c->e0, np, dc, dc + maxn, nm,
                          dc + 2 * maxn, dc + 3 * maxn, dg, 0))
                return fail(c, EDG_ERR_HIP, "continued fraction launch failed");
            HIPCHK(c, hipMemcpyAsync(g.data() + (size_t)(ispin - 1) * 2 * nw, dg, sizeof(double) * 2 * nw,
                                     hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
        // G_cl assembly (dmft_green.f90:111-118)
        for (int iw = 0; iw < nw; ++iw) {
            const double r1 = g[2 * iw], i1 = g[2 * iw + 1];
            const double r2 = g[2 * nw + 2 * iw], i2 = g[2 * nw + 2 * iw + 1];
            double *o1 = G_cl + 2 * ((size_t)(iorb - 1) * nw + iw);
            if (nspin == 1 && even) {
                o1[0] = r1;
                o1[1] = i1;
            } else if (nspin == 1) {
                o1[0] = 0.5 * (r1 + r2);
                o1[1] = 0.5 * (i1 + i2);
            } else {
                o1[0] = r1;
                o1[1] = i1;
                double *o2 = G_cl + 2 * ((size_t)norb * nw + (size_t)(iorb - 1) * nw + iw);
                o2[0] = r2;
                o2[1] = i2;
            }
        }
        if (coeffs) {
            for (int ispin = 1; ispin <= 2; ++ispin)
                for (int q = 0; q < 4; ++q)
                    std::memcpy(coeffs + (((size_t)(iorb - 1) * 2 + (ispin - 1)) * 4 + q) * maxn,
                                keep[ispin - 1].data() + q * maxn, sizeof(double) * maxn);
        }
    }
    HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_gf_ev = true;
    return 0;
}

template <typename T>
int hx_t(edg_ctx *c, int nu, int nd, const double *x, double *y, int py)
{
    HvOp op;
    if (int rc = make_op(c, nu, nd, &op)) return rc;
    const long long D = (long long)op.nup * op.ndn, Dp = round_up(D, kPad);
    if (int rc = ensure(c, c->tmp, 2 * Dp * sizeof(T))) return rc;
    T *dx = (T *)c->tmp.p, *dy = dx + Dp;
    if (int rc = upload<T>(c, dx, x, D)) return rc;
    if (py)
        if (int rc = upload<T>(c, dy, y, D)) return rc;
    HvArgs a = base_args(c);
    a.x = dx;
    a.y = dy;
    a.accumulate = py;
    if (launch_hv<T>(c->stream, op, HV_PLAIN, c->cols, c->acc64, a) < 0) return fail(c, EDG_ERR_HIP, "H.v launch failed");
    ++c->n_hv;
    return download<T>(c, y, dy, D);
}

template <typename T>
int lanczos_t(edg_ctx *c, int nu, int nd, const double *v_init, int *nstep, double *a, double *b)
{
    HvOp op;
    if (int rc = make_op(c, nu, nd, &op)) return rc;
    const long long D = (long long)op.nup * op.ndn, Dp = round_up(D, kPad);
    const int nring = c->fused ? 5 : 3;
    if (int rc = admit(c, nring * Dp * sizeof(T))) return rc;
    if (int rc = ensure(c, c->ring, nring * Dp * sizeof(T))) return rc;
    if (int rc = upload<T>(c, (T *)c->ring.p, v_init, D)) return rc;
    return run_lanczos<T>(c, lane_of(c, 0), op, (T *)c->ring.p, Dp, false, c->P.maxnstep,
                          c->fused ? (T *)c->ring.p + 3 * Dp : nullptr, nstep, a, b, nullptr, nullptr);
}

template <typename T>
int apply_c_t(edg_ctx *c, int nu, int nd, const double *vec, int pm, int iorb, int ispin, double *out)
{
    const int ns = c->P.norb + c->P.nbath;
    const int nu2 = nu + (ispin == 1 ? (pm == 1 ? 1 : -1) : 0);
    const int nd2 = nd + (ispin == 2 ? (pm == 1 ? 1 : -1) : 0);
    if (int rc = check_sector(c, nu2, nd2)) return rc;
    HvOp op0, op1;
    if (int rc = make_op(c, nu, nd, &op0)) return rc;
    if (int rc = make_op(c, nu2, nd2, &op1)) return rc;
    const long long D0 = (long long)op0.nup * op0.ndn, D1 = (long long)op1.nup * op1.ndn;
    const long long P0 = round_up(D0, kPad), P1 = round_up(D1, kPad);
    if (int rc = ensure(c, c->tmp, (P0 + P1) * sizeof(T))) return rc;
    if (int rc = ensure(c, c->mapbuf, sizeof(uint32_t) * 8192)) return rc;
    T *din = (T *)c->tmp.p, *dout = din + P0;
    if (int rc = upload<T>(c, din, vec, D0)) return rc;
    const SpinTab &tout = ispin == 1 ? op1.up : op1.dn;
    if (launch_cmap(c->stream, tout.pattern, tout.n, ns, ispin == 1 ? nu : nd, iorb - 1, pm,
                    (uint32_t *)c->mapbuf.p))
        return fail(c, EDG_ERR_HIP, "cmap launch failed");
    if (launch_apply_c<T>(c->stream, din, dout, op1.nup, op1.ndn, op0.nup, ispin, (const uint32_t *)c->mapbuf.p,
                          ispin == 2 ? (nu & 1) : 0))
        return fail(c, EDG_ERR_HIP, "apply_c launch failed");
    return download<T>(c, out, dout, D1);
}

template <typename T>
int time_hv_t(edg_ctx *c, int nu, int nd, int warmup, int reps, double *ms, double *bytes)
{
    HvOp op;
    if (int rc = make_op(c, nu, nd, &op)) return rc;
    const long long D = (long long)op.nup * op.ndn, Dp = round_up(D, kPad);
    if (int rc = ensure(c, c->tmp, 2 * Dp * sizeof(T))) return rc;
    T *dx = (T *)c->tmp.p, *dy = dx + Dp;
    c->have_gs_ev = false;
    if (launch_start<T>(c->stream, dx, D, 12345)) return fail(c, EDG_ERR_HIP, "start launch failed");
    HvArgs a = base_args(c);
    a.x = dx;
    a.y = dy;
    for (int i = 0; i < warmup; ++i)
        if (launch_hv<T>(c->stream, op, HV_PLAIN, c->cols, c->acc64, a) < 0) return fail(c, EDG_ERR_HIP, "H.v launch failed");
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    for (int i = 0; i < reps; ++i)
        if (launch_hv<T>(c->stream, op, HV_PLAIN, c->cols, c->acc64, a) < 0) return fail(c, EDG_ERR_HIP, "H.v launch failed");
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev[1]));
    float t = 0;
    HIPCHK(c, hipEventElapsedTime(&t, c->ev[0], c->ev[1]));
    *ms = (double)t / reps;
    *bytes = 2.0 * sizeof(T) * (double)D;
    return 0;
}

}  // namespace

// ============================================================== C ABI

extern "C" {

void edg_default_params(edg_params *p)
{
    std::memset(p, 0, sizeof(*p));
    p->norb = 1;
    p->nbath = 5;
    p->nspin = 1;
    p->U = 1.0;
    p->J = 0.3;
    p->Up = -1.0;
    p->Jp = -1.0;
    p->mu = 0.5;
    p->maxnstep = 80;
    p->diag_solver = 1;
    p->precision = EDG_FP64;
    p->both_spins = 0;
    p->seed = 1;
}

int edg_device_count(int *n)
{
    int k = 0;
    if (hipGetDeviceCount(&k) != hipSuccess) {
        (void)hipGetLastError();
        k = 0;
    }
    *n = k;
    return 0;
}

int edg_create(const edg_params *p, int device, edg_ctx **out)
{
    if (!p || !out) return EDG_ERR_ARG;
    *out = nullptr;
    if (p->diag_solver != 1) return EDG_ERR_ARG;  // ARPACK path (arpack_solver.f90) not provided
    if (p->norb < 1 || p->nbath < 0 || p->norb + p->nbath > 15 || p->norb > kMaxNorb ||
        p->norb * p->nbath > kMaxHops || p->maxnstep < 1 || p->maxnstep > 1000 ||
        p->maxnstep_green < 0 || p->maxnstep_green > 1000 ||
        (p->nspin != 1 && p->nspin != 2) || (p->precision != EDG_FP32 && p->precision != EDG_FP64))
        return EDG_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        (void)hipGetLastError();
        return EDG_ERR_HIP;
    }
    edg_ctx *c = new edg_ctx();
    c->P = *p;
    if (c->P.Up < 0) c->P.Up = c->P.U - 2.0 * c->P.J;  // dmft_params.f90:90-92
    if (c->P.Jp < 0) c->P.Jp = c->P.J;                  // dmft_params.f90:94-96
    if (c->P.maxnstep_green <= 0) c->P.maxnstep_green = c->P.maxnstep;
    c->device = device;
    default_schedule(p->norb, p->precision == EDG_FP32, &c->cols, &c->acc64);
    auto bad = [&]() {
        (void)hipGetLastError();
        delete c;
        return EDG_ERR_HIP;
    };
    if (hipSetDevice(device) != hipSuccess) return bad();
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bad();
    if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return bad();
    if (!dev_malloc((void **)&c->partial, sizeof(double) * 16 * 8192 + 64, device)) return bad();
    const int mx = (p->maxnstep > p->maxnstep_green ? p->maxnstep : p->maxnstep_green) + 8;
    const size_t fco = ((sizeof(double) * (3 * mx + 40) + 255) / 256) * 256;  // FusedCoef table offset
    if (!dev_malloc(&c->scal_mem, fco + sizeof(FusedCoef) * (mx + 1), device)) return bad();
    double *s = (double *)c->scal_mem;
    c->sc.alpha = s;
    c->sc.beta = s + mx;
    c->sc.scale = s + 2 * mx;
    c->sc.misc = s + 3 * mx;
    c->sc.flags = (int *)(s + 3 * mx + 32);
    c->sc.fcoef = (FusedCoef *)((char *)c->scal_mem + fco);
    for (auto &e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) return bad();
    // the second stream of the Green's-function chains (default priority)
    if (hipStreamCreateWithFlags(&c->s2, hipStreamNonBlocking) != hipSuccess) return bad();
    if (!dev_malloc((void **)&c->partial2, sizeof(double) * 16 * 8192 + 64, device)) return bad();
    if (!dev_malloc(&c->scal_mem2, fco + sizeof(FusedCoef) * (mx + 1), device)) return bad();
    double *s2 = (double *)c->scal_mem2;
    c->sc2.alpha = s2;
    c->sc2.beta = s2 + mx;
    c->sc2.scale = s2 + 2 * mx;
    c->sc2.misc = s2 + 3 * mx;
    c->sc2.flags = (int *)(s2 + 3 * mx + 32);
    c->sc2.fcoef = (FusedCoef *)((char *)c->scal_mem2 + fco);
    for (auto &e : c->ev2)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return bad();
    *out = c;
    return EDG_OK;
}

int edg_destroy(edg_ctx *c)
{
    if (!c) return EDG_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->s2) (void)hipStreamSynchronize(c->s2);
    for (auto &kv : c->tabs)
        for (void *p : kv.second.allocs) (void)hipFree(p);
    DevBuf *bufs[] = {&c->dparams, &c->basis, &c->ring, &c->gs[0], &c->gs[1], &c->stage, &c->tmp,
                      &c->coefbuf, &c->gbuf, &c->omegabuf, &c->mapbuf, &c->ring2, &c->mapbuf2, &c->sigbuf,
                      &c->robuf, &c->robuf2};
    for (DevBuf *b : bufs)
        if (b->p) (void)hipFree(b->p);
    if (c->partial) (void)hipFree(c->partial);
    if (c->scal_mem) (void)hipFree(c->scal_mem);
    if (c->partial2) (void)hipFree(c->partial2);
    if (c->scal_mem2) (void)hipFree(c->scal_mem2);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : c->ev2)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->s2) (void)hipStreamDestroy(c->s2);
    delete c;
    return EDG_OK;
}

const char *edg_last_error(const edg_ctx *c) { return c ? c->err.c_str() : noalloc_last_error(); }

int edg_set_memory_budget(edg_ctx *c, unsigned long long bytes)
{
    if (!c) return EDG_ERR_ARG;
    c->budget = bytes;
    return EDG_OK;
}

int edg_set_tuning(edg_ctx *c, int cols_per_block, int fp64_accumulate)
{
    if (!c || (cols_per_block != 1 && cols_per_block != 2 && cols_per_block != 4 && cols_per_block != 8 &&
               cols_per_block != 16 && cols_per_block != 32 && cols_per_block != 64))
        return EDG_ERR_ARG;
    c->cols = cols_per_block;
    c->acc64 = fp64_accumulate != 0;
    return EDG_OK;
}

int edg_set_reorthogonalization(edg_ctx *c, int mode)
{
    if (!c || mode < 0 || mode > 2) return EDG_ERR_ARG;
    c->reorth = mode;
    return EDG_OK;
}

int edg_reorth_stats(edg_ctx *c, long long *events, long long *steps)
{
    if (!c) return EDG_ERR_ARG;
    if (events) *events = c->ro_events;
    if (steps) *steps = c->ro_steps;
    return EDG_OK;
}

int edg_set_fused_steps(edg_ctx *c, int on)
{
    if (!c) return EDG_ERR_ARG;
    c->fused = on != 0;
    return EDG_OK;
}

int edg_spin_table(edg_ctx *c, const double *em, const double *ek, const double *vmk, int ispin,
                   int ne, int *n, int *kmax, unsigned *pattern, int *hop_col, double *hop_val,
                   int *hop_cnt, double *energy)
{
    if (!c) return EDG_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (ispin != 1 && ispin != 2) return fail(c, EDG_ERR_ARG, "ispin must be 1 or 2");
    if (int rc = check_sector(c, ne, 0)) return rc;
    if (int rc = set_params(c, em, ek, vmk)) return rc;
    const SpinTab *t;
    if (int rc = get_tab(c, ispin, ne, &t)) return rc;
    *n = t->n;
    *kmax = c->P.norb * c->P.nbath;  // slots beyond norb*nbath are padding only
    const size_t N = t->n, K = t->K, Kr = (size_t)c->P.norb * c->P.nbath;
    if (pattern) HIPCHK(c, hipMemcpyAsync(pattern, t->pattern, N * 4, hipMemcpyDeviceToHost, c->stream));
    std::vector<HopD> hd(K * N);
    std::vector<uint8_t> cnt(N);
    if (K) HIPCHK(c, hipMemcpyAsync(hd.data(), t->hd, K * N * sizeof(HopD), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(cnt.data(), t->cnt, N, hipMemcpyDeviceToHost, c->stream));
    if (energy) HIPCHK(c, hipMemcpyAsync(energy, t->energy, N * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t k = 0; k < Kr; ++k)
        for (size_t r = 0; r < N; ++r) {
            const bool used = k < cnt[r];
            if (hop_col) hop_col[k * N + r] = used ? (int)hd[r * K + k].col : -1;
            if (hop_val) hop_val[k * N + r] = used ? hd[r * K + k].val : 0.0;
        }
    if (hop_cnt)
        for (size_t i = 0; i < N; ++i) hop_cnt[i] = cnt[i];
    return EDG_OK;
}


int edg_hx(edg_ctx *c, const double *em, const double *ek, const double *vmk, int ne_up, int ne_dn,
           const double *x, double *y, int py)
{
    if (!c || !x || !y) return EDG_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (int rc = set_params(c, em, ek, vmk)) return rc;
    return c->P.precision == EDG_FP32 ? hx_t<float>(c, ne_up, ne_dn, x, y, py)
                                      : hx_t<double>(c, ne_up, ne_dn, x, y, py);
}


int edg_lanczos_iteration(edg_ctx *c, const double *em, const double *ek, const double *vmk,
                          int ne_up, int ne_dn, const double *v_init, int *nstep, double *a, double *b)
{
    if (!c || !v_init || !nstep || !a || !b) return EDG_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (int rc = set_params(c, em, ek, vmk)) return rc;
    return c->P.precision == EDG_FP32 ? lanczos_t<float>(c, ne_up, ne_dn, v_init, nstep, a, b)
                                      : lanczos_t<double>(c, ne_up, ne_dn, v_init, nstep, a, b);
}

int edg_ground_state(edg_ctx *c, const double *em, const double *ek, const double *vmk, int nsector,
                     const int *sectors, int *isector, double *e0)
{
    if (!c || nsector < 1 || !sectors || !isector || !e0) return EDG_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (int rc = set_params(c, em, ek, vmk)) return rc;
    c->n_hv = c->n_steps = 0;
    return c->P.precision == EDG_FP32 ? ground_state_t<float>(c, nsector, sectors, isector, e0)
                                      : ground_state_t<double>(c, nsector, sectors, isector, e0);
}

int edg_gs_info(edg_ctx *c, int *ne_up, int *ne_dn, long long *dim, double *e0_ritz, double *e0,
                double *residual, int *nstep)
{
    if (!c) return EDG_ERR_ARG;
    if (!c->has_gs) return fail(c, EDG_ERR_STATE, "no ground state held");
    if (ne_up) *ne_up = c->gs_nu;
    if (ne_dn) *ne_dn = c->gs_nd;
    if (dim) *dim = c->gs_dim;
    if (e0_ritz) *e0_ritz = c->e0_ritz;
    if (e0) *e0 = c->e0;
    if (residual) *residual = c->residual;
    if (nstep) *nstep = c->gs_nstep;
    return EDG_OK;
}

int edg_get_gs_vector(edg_ctx *c, double *vec, long long dim)
{
    if (!c || !vec) return EDG_ERR_ARG;
    if (!c->has_gs) return fail(c, EDG_ERR_STATE, "no ground state held");
    if (dim != c->gs_dim) return fail(c, EDG_ERR_ARG, "dimension mismatch");
    (void)hipSetDevice(c->device);
    return c->P.precision == EDG_FP32 ? download<float>(c, vec, (const float *)c->gs[c->gs_slot].p, dim)
                                      : download<double>(c, vec, (const double *)c->gs[c->gs_slot].p, dim);
}

int edg_set_gs(edg_ctx *c, int ne_up, int ne_dn, double e0, const double *vec, long long dim)
{
    if (!c || !vec) return EDG_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (int rc = check_sector(c, ne_up, ne_dn)) return rc;
    const int ns = c->P.norb + c->P.nbath;
    if (dim != binom(ns, ne_up) * binom(ns, ne_dn)) return fail(c, EDG_ERR_ARG, "dimension mismatch");
    const size_t es = elem_size(c);
    if (int rc = ensure(c, c->gs[c->gs_slot], round_up(dim, kPad) * es)) return rc;
    int rc = c->P.precision == EDG_FP32 ? upload<float>(c, (float *)c->gs[c->gs_slot].p, vec, dim)
                                        : upload<double>(c, (double *)c->gs[c->gs_slot].p, vec, dim);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->has_gs = true;
    c->gs_nu = ne_up;
    c->gs_nd = ne_dn;
    c->gs_dim = dim;
    c->e0 = c->e0_ritz = e0;
    c->residual = 0.0;
    c->gs_nstep = 0;
    c->gs_isector = 1;
    return EDG_OK;
}

int edg_cluster_green(edg_ctx *c, const double *em, const double *ek, const double *vmk, int nw,
                      const double *omega, double *G_cl, double *coeffs)
{
    if (!c || nw < 1 || !omega || !G_cl) return EDG_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (int rc = set_params(c, em, ek, vmk)) return rc;
    return c->P.precision == EDG_FP32 ? cluster_green_t<float>(c, nw, omega, G_cl, coeffs)
                                      : cluster_green_t<double>(c, nw, omega, G_cl, coeffs);
}

int edg_self_energy(edg_ctx *c, const double *em, const double *ek, const double *vmk, int nw,
                    const double *omega, const double *G_cl, double *Sigma, double *Z)
{
    if (!c || nw < 1 || !omega || !G_cl || !Sigma) return EDG_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (int rc = set_params(c, em, ek, vmk)) return rc;
    const int norb = c->P.norb, nspin = c->P.nspin;
    const size_t ng = (size_t)nw * norb * nspin;
    // omega | G | Sigma | Z on the device
    const size_t zoff = 8 * ((size_t)nw + 4 * ng);
    if (int rc = ensure(c, c->sigbuf, zoff + 8 * (size_t)norb * nspin)) return rc;
    char *b = (char *)c->sigbuf.p;
    double *om = (double *)b, *g = om + nw, *sg = g + 2 * ng, *z = (double *)(b + zoff);
    HIPCHK(c, hipMemcpyAsync(om, omega, 8 * (size_t)nw, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(g, G_cl, 16 * ng, hipMemcpyHostToDevice, c->stream));
    if (launch_sigma(c->stream, nw, norb, c->P.nbath, nspin, om, (const double *)c->dparams.p, c->P.mu, g, sg, z))
        return fail(c, EDG_ERR_HIP, "self-energy launch failed");
    HIPCHK(c, hipMemcpyAsync(Sigma, sg, 16 * ng, hipMemcpyDeviceToHost, c->stream));
    if (Z) HIPCHK(c, hipMemcpyAsync(Z, z, 8 * (size_t)norb * nspin, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return EDG_OK;
}

int edg_observables(edg_ctx *c, double *occ, double *docc)
{
    if (!c || !occ || !docc) return EDG_ERR_ARG;
    if (!c->has_gs) return fail(c, EDG_ERR_STATE, "no ground state held");
    (void)hipSetDevice(c->device);
    HvOp op;
    if (int rc = make_op(c, c->gs_nu, c->gs_nd, &op)) return rc;
    double *out = c->sc.misc + 6;
    int rc = c->P.precision == EDG_FP32
                 ? launch_observables<float>(c->stream, (const float *)c->gs[c->gs_slot].p, op, nullptr,
                                             c->partial, out)
                 : launch_observables<double>(c->stream, (const double *)c->gs[c->gs_slot].p, op, nullptr,
                                              c->partial, out);
    if (rc) return fail(c, EDG_ERR_HIP, "observables launch failed");
    const int norb = c->P.norb;
    std::vector<double> h(3 * norb);
    HIPCHK(c, hipMemcpyAsync(h.data(), out, sizeof(double) * 3 * norb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int m = 0; m < norb; ++m) {
        occ[m] = h[m];
        occ[norb + m] = h[norb + m];
        docc[m] = h[2 * norb + m];
    }
    return EDG_OK;
}


int edg_apply_c(edg_ctx *c, int ne_up, int ne_dn, const double *vec, int pm, int iorb, int ispin,
                double *vec_out)
{
    if (!c || !vec || !vec_out || (pm != 1 && pm != 2) || (ispin != 1 && ispin != 2) || iorb < 1 ||
        iorb > (c ? c->P.norb : 0))
        return c ? fail(c, EDG_ERR_ARG, "bad apply_c arguments") : EDG_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (c->em.empty()) {  // apply_c needs only the basis: install zero parameters
        std::vector<double> z(2 * c->P.norb * (c->P.nbath + 1) + 2 * c->P.nbath, 0.0);
        if (int rc = set_params(c, z.data(), z.data(), z.data())) return rc;
    }
    return c->P.precision == EDG_FP32 ? apply_c_t<float>(c, ne_up, ne_dn, vec, pm, iorb, ispin, vec_out)
                                      : apply_c_t<double>(c, ne_up, ne_dn, vec, pm, iorb, ispin, vec_out);
}


int edg_time_hv(edg_ctx *c, const double *em, const double *ek, const double *vmk, int ne_up, int ne_dn,
                int warmup, int reps, double *ms_per_launch, double *bytes_per_launch)
{
    if (!c || reps < 1 || warmup < 0 || !ms_per_launch || !bytes_per_launch) return EDG_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (int rc = set_params(c, em, ek, vmk)) return rc;
    return c->P.precision == EDG_FP32 ? time_hv_t<float>(c, ne_up, ne_dn, warmup, reps, ms_per_launch, bytes_per_launch)
                                      : time_hv_t<double>(c, ne_up, ne_dn, warmup, reps, ms_per_launch, bytes_per_launch);
}

int edg_time_stream(edg_ctx *c, long long n, int reps, double *copy_ms, double *add_ms)
{
    if (!c || n < 4 || reps < 1 || !copy_ms || !add_ms) return EDG_ERR_ARG;
    (void)hipSetDevice(c->device);
    n &= ~3LL;
    const long long np = round_up(n, kPad);
    if (int rc = ensure(c, c->tmp, 3 * np * sizeof(float))) return rc;
    float *x = (float *)c->tmp.p, *z = x + np, *y = z + np;
    if (launch_start<float>(c->stream, x, n, 1) || launch_start<float>(c->stream, z, n, 2))
        return fail(c, EDG_ERR_HIP, "start launch failed");
    c->have_gs_ev = false;
    for (int pass = 0; pass < 2; ++pass) {
        auto one = [&]() {
            return pass == 0 ? launch_stream_copy(c->stream, x, y, n) : launch_stream_add(c->stream, x, z, y, n);
        };
        for (int i = 0; i < 5; ++i)
            if (one()) return fail(c, EDG_ERR_HIP, "stream launch failed");
        HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
        for (int i = 0; i < reps; ++i)
            if (one()) return fail(c, EDG_ERR_HIP, "stream launch failed");
        HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
        HIPCHK(c, hipEventSynchronize(c->ev[1]));
        float t = 0;
        HIPCHK(c, hipEventElapsedTime(&t, c->ev[0], c->ev[1]));
        (pass == 0 ? *copy_ms : *add_ms) = (double)t / reps;
    }
    return EDG_OK;
}

int edg_last_solve_stats(edg_ctx *c, double *phase_ms, long long *n_hv, long long *n_steps)
{
    if (!c || !phase_ms) return EDG_ERR_ARG;
    phase_ms[0] = phase_ms[1] = 0.0;
    float t;
    if (c->have_gs_ev && hipEventElapsedTime(&t, c->ev[0], c->ev[1]) == hipSuccess) phase_ms[0] = t;
    if (c->have_gf_ev && hipEventElapsedTime(&t, c->ev[2], c->ev[3]) == hipSuccess) phase_ms[1] = t;
    (void)hipGetLastError();
    if (n_hv) *n_hv = c->n_hv;
    if (n_steps) *n_steps = c->n_steps;
    return EDG_OK;
}

}  // extern "C"
