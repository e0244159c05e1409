// Batched one- and two-orbital ED solver and request loop (C ABI: edg_batch_*
// in include/edgpu.h).
//
// Each problem moves through the stages of edg_ground_state +
// edg_cluster_green:
//   QUEUED   waiting for device memory (admission by bytes)
//   GS       ground-state Lanczos chain (stored Krylov basis), batched
//   GS_READ  alpha/beta on their way to the host; then the tridiagonal
//            eigenvector, gs = sum_j c_j q_j, the FP64 Rayleigh quotient and
//            the c+/c start vectors (per-problem launches)
//   CHAINS   the c+/c Lanczos chains of every orbital and spin (ring of 3
//            vectors each), batched
//   CH_READ  chain coefficients on their way to the host; then the continued
//            fractions on the device
//   CF_READ  G(i w) on its way to the host; then DONE and memory released.
// Every Lanczos chain in flight (ground state or c+/c, any problem, any
// step) is advanced by the same launches per step (batch.hip: one batched
// H.v per launch class and the axpy; one step advance per chunk of up to 8
// steps); the stages between chains are short per-problem launches on
// side streams (problems dealt round-robin over kSides of them).
// Host-side waits use events polled without blocking, so new problems are
// admitted while others run (continuous batching).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <exception>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "../../include/edgpu.h"
#include "devmem.h"
#include "edgpu_internal.h"

namespace edg {
int tridiag_lowest(int n, const double *d, const double *e, double *lambda, double *z);
}

using namespace edg;

namespace {

constexpr long long kPadB = 64;       // element alignment of vectors
constexpr int kPartialCap = 1 << 17;  // doubles of the shared partial-sum scratch
// dynamic LDS of a batched tile: the 160 KiB of a gfx950 CU less the static
// LDS of the batched kernels (k_bhv<double, double, 16> holds 14 KiB of
// per-column neighbour offsets and elements, tables and reduction scratch)
constexpr size_t kLdsMax = 144 * 1024;
constexpr int kClasses = 9;            // H.v launch classes, chain_class()
constexpr int kMaxChains = 8;          // c+/c chains per problem: 2 orbitals x 2 spins x 2
constexpr int kChunk = 8;              // batched steps launched per chunk
// Side streams of the stage work, problems dealt round-robin: the stage
// kernels are small (a few workgroups each), and a batch whose problems all
// reach a stage together queued them on one stream one after another while
// the batched steps had no chains to run.  With the step stream this is the
// device's 4 hardware queues.
constexpr int kSides = 3;
constexpr int kGraphAfter = 2;
// host threads of the ground-state eigensolves (tridiagonal of up to 1000
// steps, ~30 us at 200), at least kEigenPerThread problems each
constexpr int kEigenThreads = 8;
constexpr int kEigenPerThread = 8;         // chunks of an unchanged chain set before one is captured
constexpr size_t kDescMin = 1 << 20;   // initial bytes of the descriptor buffer
constexpr int kAhead = 2;              // chunks the host may run ahead of the device
constexpr int kChunkEvents = 4;        // > kAhead
// Problems whose state vector takes kLargeVec bytes or more are admitted at
// most kMaxLarge at a time: many such vectors in flight compete for the MALL
// (a batch of 8 Ns=14 problems, 47 MB FP32 vectors, measured 41.5 ms per
// problem admitted one at a time against 44.8 all at once), while two let
// one problem's ground-state chain run beside the other's c+/c chains
// (four Ns=14 problems 18.1-18.2 -> 18.6-18.9 solves/s against one at a time)
constexpr size_t kLargeVec = size_t(32) << 20;
constexpr int kMaxLarge = 2;
// Chains of kDirectMin elements or more run their steps as launches of the
// single-problem kernels (edg_cluster_green's, arguments from the host)
// instead of the batched ones: one chain's H.v then fills the GPU, and the
// batched tile's descriptor reads (chain index, step and stop flag before
// any load of the vectors) cost 25-30% per step at Ns=14.
constexpr long long kDirectMin = 4ll << 20;

long long rup(long long n, long long m) { return (n + m - 1) / m * m; }

long long binom_ll(int n, int k)
{
    if (k < 0 || k > n) return 0;
    long long r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

// Largest hop count of one spin half with ne electrons on norb impurity +
// nbath bath sites: a configuration with a occupied impurity orbitals has
// a (nbath - (ne - a)) + (norb - a)(ne - a) pairs of unlike occupation
// (zero hybridizations only lower it).
// Exact largest hop count of one spin half over its ne-electron patterns
// (impurity orbital m = bit m, bath site l = bit norb + l): pairs (m, l)
// with V_ml != 0 and unlike occupation.  At most 2^15 patterns.
int kmax_exact(int norb, int nbath, int ne, const double *v_s)
{
    const int ns = norb + nbath;
    int best = 0;
    for (uint32_t p = 0; p < (1u << ns); ++p) {
        if (__builtin_popcount(p) != ne) continue;
        int k = 0;
        for (int m = 0; m < norb; ++m)
            for (int l = 0; l < nbath; ++l)
                if (v_s[l * norb + m] != 0.0 && (((p >> m) ^ (p >> (norb + l))) & 1u)) ++k;
        best = std::max(best, k);
    }
    return best;
}

int kmax_hops(int norb, int nbath, int ne)
{
    int k = 0;
    for (int a = 0; a <= norb; ++a) {
        const int bo = ne - a;
        if (bo < 0 || bo > nbath) continue;
        k = std::max(k, a * (nbath - bo) + (norb - a) * bo);
    }
    return k;
}

enum Stage { ST_QUEUED, ST_GS, ST_GS_READ, ST_CHAINS, ST_CH_READ, ST_CF_READ, ST_DONE };

struct Chain {
    BInst bi{};
    bool direct = false;  // steps as single-problem launches (D >= kDirectMin)
    int ax_n = 0;         // partial sums of the last axpy of a direct chain
    int launched = 0;     // steps launched
    bool exists = false;
    bool pending = false;  // started on the side stream, not yet in the batched steps
    bool active = false;
    int iorb = 0, ispin = 0, pm = 0;
    long long D = 0;
};

// Bump allocator over one device block per problem.  A request past the
// block's end is refused (returns the block start and sets overflow), and
// callers check overflow before any kernel uses what they took.
struct Arena {
    char *base = nullptr;
    size_t used = 0, cap = 0;
    bool overflow = false;
    void *take(size_t bytes)
    {
        const size_t off = (used + 255) & ~size_t(255);
        if (off + bytes > cap) {
            overflow = true;
            return base;
        }
        used = off + bytes;
        return base + off;
    }
};

// Pinned host memory is taken in slabs and cut into the problems' areas;
// a destroyed queue leaves its slabs to later queues of the process
// (pinning pages costs far more than the few transfers of a small problem).
struct PinSlab {
    int device;
    void *p;
    size_t bytes;
};
constexpr size_t kPinSlab = 16u << 20;
constexpr size_t kPinCacheMax = 1ull << 30;
std::mutex g_pin_mu;
std::vector<PinSlab> g_pin_cache;
size_t g_pin_cached = 0;

struct Job {
    long long id = 0;
    int lane = 0;  // its side stream (stage work) and that stream's partial-sum scratch
    edg_problem p{};
    std::vector<double> ek, vmk;
    Stage st = ST_QUEUED;
    size_t bytes = 0;
    bool large = false;  // state vector of kLargeVec bytes or more
    Arena ar;
    int norb = 1;
    double *dparams = nullptr;  // em(norb,2) | ek(nbath,2) | vmk(norb,nbath,2)
    SpinTab tab[2][3];          // [spin][ne - N_spin + 1]
    bool has_tab[2][3] = {};
    // c+ maps of the N + 1 tables, c maps of the N - 1 tables ([orbital][rank])
    uint32_t *tmap[2][3] = {};
    void *gs = nullptr;
    double *misc = nullptr;     // [1..3] RQ sums, [4] |combine|, [5] 1/|combine|, [8+k] |v_init| of chain k
    double *coef = nullptr;     // device coefficient staging
    double *gdev = nullptr;     // norb x 2 spins x 2 nw
    Chain gsc;
    Chain ch[kMaxChains];       // chain 4 (iorb - 1) + k: (up,+), (up,-), (dn,+), (dn,-)
    double *host = nullptr;     // pinned readback / upload area
    hipEvent_t ev = nullptr;
    // recorded on the side stream after the stage work that started chains;
    // the chains join the batched steps once the host sees it complete
    hipEvent_t start_ev = nullptr;
    bool start_pending = false;
    double e0 = 0.0;
    int gs_nstep = 0;
    std::vector<double> G;
    std::vector<double> coeffs;  // (mg, 4, 2, norb) column-major: ap bp am bm per spin and orbital
};

}  // namespace

struct edg_batch {
    int device = 0;
    // stream: the batched Lanczos steps (and the reads of finished chains);
    // side[lane]: per-problem stage work (tables, start vectors, combine, RQ,
    // apply_c, continued fractions) of the problems of that lane, which
    // overlaps the steps and the other lanes' stages.  Stage streams never
    // wait on each other: the host moves a problem to its next stage, and
    // lets its new chains join the steps, only after an event shows the
    // previous stage's work is done.  The one device-side wait is the fork
    // and join around the direct chains' steps that run on the last side
    // stream (run_direct).  The host stays at most kAhead chunks of kChunk
    // steps ahead of the device, so chains that become ready join within a
    // few steps.
    hipStream_t stream = nullptr;
    hipStream_t side[kSides] = {};
    double *side_part[kSides] = {};  // partial-sum scratch of each side stream's stage kernels
    int next_lane = 0;
    hipEvent_t chunk_ev[4] = {};
    // direct chains split over the step stream and the last side stream
    // within a chunk: fork (step stream -> side) and join (side -> step)
    hipEvent_t dfork = nullptr, djoin = nullptr;
    int chunk_head = 0, chunk_n = 0;
    bool fp64 = false;
    size_t es = 4;
    int nspin = 1;
    int nw = 0;
    int ncu = 0;
    double *omega = nullptr;
    unsigned long long budget = 0, in_use = 0;
    DevHeap *heap = nullptr;  // device memory of the problems in flight
    std::string err;
    std::deque<Job *> queue;
    std::vector<Job *> running;
    std::deque<Job *> finished;
    long long next_id = 1;
    // chains in the batched launches, and the direct ones
    std::vector<Chain *> active, direct;
    bool dirty = true;
    bool moved = false;  // a direct chain joined or left
    char *ddesc = nullptr;
    size_t ddesc_cap = 0;
    char *hdesc[2] = {nullptr, nullptr};
    size_t hdesc_cap[2] = {0, 0};
    hipEvent_t hdesc_ev[2] = {nullptr, nullptr};
    bool hdesc_used[2] = {false, false};
    int flip = 0;
    // chains in launch classes {short, long columns} x {8, 16 hop slots}
    const BInst *d_inst = nullptr;
    const int *d_axoff = nullptr, *d_axown = nullptr;
    const int *d_hvoff[kClasses] = {}, *d_hvown[kClasses] = {};
    int cls_start[kClasses] = {}, cls_n[kClasses] = {};
    int cls_tiles[kClasses] = {};
    size_t cls_lds[kClasses] = {};  // bytes
    int nax = 0;
    bool use_graphs = true;
    hipGraphExec_t gexec[kChunk + 1] = {};  // [k]: a chunk of k steps of the current chain set
    // replaced graphs may still have launches queued: destroyed once the
    // stream has passed the event recorded after their last launch
    std::vector<std::pair<hipGraphExec_t, hipEvent_t>> retired;
    int since_rebuild = 0;  // chunks launched since the chain set last changed
    std::vector<double *> pinned_free;  // problem areas of pinned_n doubles, cut from slabs
    std::vector<PinSlab> slabs;
    size_t pinned_n = 0;
    std::vector<hipEvent_t> ev_free;  // events of finished problems, for the next ones
    double st_steps = 0, st_chain_steps = 0, st_done = 0, st_graph = 0, st_rebuild = 0;
    long long st_inst = 0, st_exec_destroyed = 0;  // graph instantiations / execs destroyed
    // host seconds inside edg_batch_step: admission (tables, start vectors),
    // stage transitions (progress: read-backs, eigenvectors, combine, apply_c,
    // continued fractions), queueing the steps, and blocked on events
    double st_t_admit = 0, st_t_progress = 0, st_t_steps = 0, st_t_wait = 0;
    // H.v tile schedule by size class: the first entry with nsite <= nsite_max
    // applies to every chain of a problem
    struct Sched {
        int nsite_max, tile, wide;
    };
    std::vector<Sched> sched{{64, kBatchTile, -1}};  // wide = -1: nup >= kBatchWide
};

namespace {

int fail(edg_batch *b, int code, const std::string &msg)
{
    b->err = msg;
    return code;
}

#define BCHK(b, call)                                                                             \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return fail(b, EDG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

int norb_of(const edg_problem &p) { return p.norb >= 2 ? 2 : 1; }
int nsite_of(const edg_problem &p) { return norb_of(p) + p.nbath; }
int green_len(const edg_problem &p) { return p.maxnstep_green > 0 ? p.maxnstep_green : p.maxnstep; }

// Chains a problem runs: spin-down chains only when spin symmetry does not
// give them (dmft_green.f90:76-80) or when both_spins is set.
bool runs_spin2(const edg_batch *b, const edg_problem &p)
{
    const bool even = !(b->nspin == 2 || p.nup != p.ndn);
    return !even || p.both_spins;
}

size_t tab_bytes(int n, int norb, int nbath)
{
    const size_t K = std::max(norb * nbath, kMinSlots);
    const size_t orb = norb >= 2 ? 4 * (size_t)norb * norb * n + 256 : 0;
    return 4 * (size_t)n + K * n * (sizeof(HopF) + sizeof(HopD)) + n + 4 + 8 * (size_t)n + 4 * (size_t)n +
           n + 16 * (size_t)n * (2 + 1 + 4) + orb + 4 * (size_t)norb * n + 17 * 256;
}

// A chain's device state block: step index (int) at byte 0, flags (two
// ints: stop, number of steps) at byte 8, then alpha[maxn], beta[maxn + 1]
// and scale[maxn + 1] from byte 256 (kHead doubles).  A read-back copies the
// block's first chain_host_n(maxn) doubles in one transfer.
constexpr int kHead = 32;
size_t chain_host_n(int m) { return kHead + 3 * (size_t)m + 2; }

// Pinned host area of a problem (doubles): the misc sums, the state blocks
// of the chains read back, the coefficient staging and G.
size_t host_chain(int k, int mg) { return 16 + (size_t)k * chain_host_n(mg); }
size_t host_coef(int mg) { return host_chain(kMaxChains, mg); }
size_t host_g(int mg) { return host_coef(mg) + 16 * (size_t)mg; }  // (ap bp am bm) x 2 spins x 2 orbitals
constexpr int kMaxLen = 1000;  // longest Lanczos run of a problem

size_t state_bytes(int maxn, int hvn, int axn)
{
    return 8 * kHead + ((size_t)maxn + 1) * 3 * 8 + 16 + (size_t)(hvn + axn) * 8;
}

// Tile shape of a chain with columns of nup rows under the engine's
// schedule: long-column tiles (kBatchWideCols columns per 512-thread
// workgroup when they fit LDS, else one) or cb = clamp(tile / nup, 1,
// kBatchMaxCols) columns per 256-thread workgroup.
struct Shape {
    int wide, cb;
};
int wide_cols(const edg_batch *b, long long nup)
{
    return (size_t)kBatchWideCols * nup * b->es <= kLdsMax ? kBatchWideCols : 1;
}
// Two-orbital chains: the single-problem default of kBatchKanColsF32
// columns per tile with FP32 vectors (FP64 accumulation), one with FP64.
int kan_cols(const edg_batch *b) { return b->fp64 ? 1 : kBatchKanColsF32; }
Shape chain_shape(const edg_batch *b, int nsite, long long nup, int norb = 1)
{
    if (norb >= 2) return {1, kan_cols(b)};
    for (const auto &e : b->sched)
        if (nsite <= e.nsite_max) {
            const int wide = e.wide < 0 ? (nup >= kBatchWide) : e.wide;
            int cb = e.tile / (int)std::max<long long>(nup, 1);
            cb = std::min(std::max(cb, 1), kBatchMaxCols);
            return {wide, wide ? wide_cols(b, nup) : cb};
        }
    const int wide = nup >= kBatchWide ? 1 : 0;
    return {wide, wide ? wide_cols(b, nup) : 1};
}

int chain_hvn(const edg_batch *b, int nsite, long long nup, long long ndn, int norb = 1)
{
    const Shape sh = chain_shape(b, nsite, nup, norb);
    return (int)((ndn + sh.cb - 1) / sh.cb);
}

// Axpy workgroups of a chain: 32 elements per thread.  Each workgroup first
// reduces the chain's H.v partials, and every H.v tile of the next step
// reduces the chain's axpy partials, so fewer, fuller workgroups cut both.
int chain_axn(long long D)
{
    long long g = (D + kBlock * 32 - 1) / (kBlock * 32);
    return (int)std::min<long long>(std::max<long long>(g, 1), 1024);
}

// Partial-sum slots of a chain: the batched tiles and axpy workgroups, or
// for a direct chain bounds of the single-problem kernels' grids (one
// workgroup per sector column at most, kAxpyMaxGrid axpy workgroups).
int chain_hv_parts(const edg_batch *b, int nsite, long long nup, long long ndn, int norb)
{
    return nup * ndn >= kDirectMin ? (int)ndn : chain_hvn(b, nsite, nup, ndn, norb);
}
int chain_ax_parts(long long D) { return D >= kDirectMin ? std::max(chain_axn(D), kAxpyMaxGrid) : chain_axn(D); }

// Device bytes of a problem in flight.
size_t problem_bytes(const edg_batch *b, const edg_problem &p)
{
    const int ns = nsite_of(p), no = norb_of(p);
    size_t bytes = 64 * 1024 + 8 * (size_t)(2 * no + 2 * p.nbath + 2 * no * p.nbath);
    for (int s = 0; s < 2; ++s) {
        const int n0 = s == 0 ? p.nup : p.ndn;
        for (int d = -1; d <= 1; ++d)
            if (n0 + d >= 0 && n0 + d <= ns) bytes += tab_bytes((int)binom_ll(ns, n0 + d), no, p.nbath);
    }
    const long long nu = binom_ll(ns, p.nup), nd = binom_ll(ns, p.ndn);
    const long long D = nu * nd, Dp = rup(D, kPadB);
    bytes += ((size_t)p.maxnstep + 1) * Dp * b->es + Dp * b->es;
    bytes += state_bytes(p.maxnstep, chain_hv_parts(b, ns, nu, nd, no), chain_ax_parts(D));
    const int mg = green_len(p);
    bytes += 64 * 8 + 8 * (8 * (size_t)no * std::max(mg, p.maxnstep) + 64) + 8 * 4 * (size_t)no * b->nw;
    for (int k = 0; k < 4 * no && !p.gs_only; ++k) {
        const int ispin = (k & 3) < 2 ? 1 : 2, pm = (k & 1) ? 2 : 1;
        const int cu = p.nup + (ispin == 1 ? (pm == 1 ? 1 : -1) : 0);
        const int cd = p.ndn + (ispin == 2 ? (pm == 1 ? 1 : -1) : 0);
        if (cu < 0 || cd < 0 || cu > ns || cd > ns) continue;
        const long long cnu = binom_ll(ns, cu), cnd = binom_ll(ns, cd), Dc = cnu * cnd;
        bytes += 3 * rup(Dc, kPadB) * b->es;
        bytes += state_bytes(mg, chain_hv_parts(b, ns, cnu, cnd, no), chain_ax_parts(Dc));
    }
    return bytes + 64 * 256;
}

int validate(const edg_batch *b, const edg_problem &p)
{
    const int ns = nsite_of(p), no = norb_of(p);
    if (p.norb < 0 || p.norb > 2) return fail(const_cast<edg_batch *>(b), EDG_ERR_ARG, "norb must be 0, 1 or 2");
    if (p.nbath < 1 || ns > 15)
        return fail(const_cast<edg_batch *>(b), EDG_ERR_ARG, "nbath must be 1..15 - norb");
    if (!p.ek || !p.vmk) return fail(const_cast<edg_batch *>(b), EDG_ERR_ARG, "ek / vmk missing");
    if (p.nup < 0 || p.ndn < 0 || p.nup > ns || p.ndn > ns)
        return fail(const_cast<edg_batch *>(b), EDG_ERR_ARG, "sector outside 0..nsite");
    if (p.gs_only != 0 && p.gs_only != 1) return fail(const_cast<edg_batch *>(b), EDG_ERR_ARG, "gs_only must be 0 or 1");
    if (p.maxnstep < 1 || p.maxnstep > kMaxLen || green_len(p) > kMaxLen)
        return fail(const_cast<edg_batch *>(b), EDG_ERR_ARG, "Lanczos lengths must be 1..1000");
    for (int s = 0; s < 2; ++s) {
        const int n0 = s == 0 ? p.nup : p.ndn;
        for (int d = -1; d <= 1; ++d) {
            const int ne = n0 + d;
            if (ne < 0 || ne > ns) continue;
            if (kmax_hops(no, p.nbath, ne) > 16)
                return fail(const_cast<edg_batch *>(b), EDG_ERR_ARG, "more than 16 hops per configuration");
            const long long n = binom_ll(ns, ne);
            const Shape sh = chain_shape(b, ns, n, no);
            const size_t lds = no >= 2 ? hv2_lds((int)n, sh.cb, b->es) : (size_t)sh.cb * n * b->es;
            if (lds > kLdsMax) return fail(const_cast<edg_batch *>(b), EDG_ERR_ARG, "sector column exceeds LDS");
        }
    }
    return 0;
}

double *pinned_get(edg_batch *b)
{
    if (!b->pinned_free.empty()) {
        double *p = b->pinned_free.back();
        b->pinned_free.pop_back();
        return p;
    }
    const size_t per = (b->pinned_n * sizeof(double) + 255) & ~size_t(255);
    PinSlab sl{b->device, nullptr, 0};
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        for (size_t i = 0; i < g_pin_cache.size(); ++i)
            if (g_pin_cache[i].device == b->device && g_pin_cache[i].bytes >= per) {
                sl = g_pin_cache[i];
                g_pin_cache.erase(g_pin_cache.begin() + i);
                g_pin_cached -= sl.bytes;
                break;
            }
    }
    if (!sl.p) {
        sl.bytes = std::max(kPinSlab, per);
        if (hipHostMalloc(&sl.p, sl.bytes, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    b->slabs.push_back(sl);
    for (size_t off = sl.bytes / per * per; off >= per; off -= per)
        b->pinned_free.push_back((double *)((char *)sl.p + off - per));
    double *p = b->pinned_free.back();
    b->pinned_free.pop_back();
    return p;
}

// U' and J' of a two-orbital problem (dmft_params.f90:90-96 defaults).
double kan_up(const edg_problem &p) { return p.Up < 0 ? p.U - 2.0 * p.J : p.Up; }
double kan_jp(const edg_problem &p) { return p.Jp < 0 ? p.J : p.Jp; }

// Per-spin tables of one (spin, N) half; kc = hop slots per row, the same
// for every table of the problem so that its chains share one launch class;
// km = the table's largest hop count; pm = 1 / 2: also the c+ / c maps of
// the sector (into *map).
int build_table(edg_batch *b, Job *J, int spin, int ne, SpinTab &t, int kc, int km, int pm, uint32_t **map)
{
    const int ns = nsite_of(J->p), nbath = J->p.nbath, norb = J->norb;
    t = SpinTab{};
    t.n = (int)binom_ll(ns, ne);
    t.K = std::max(norb * nbath, kMinSlots);
    const size_t n = t.n, K = t.K;
    t.pattern = (uint32_t *)J->ar.take(4 * n);
    t.hf = (HopF *)J->ar.take(K * n * sizeof(HopF));
    t.hd = (HopD *)J->ar.take(K * n * sizeof(HopD));
    t.cnt = (uint8_t *)J->ar.take(n);
    t.energy = (double *)J->ar.take(8 * n);
    t.energyf = (float *)J->ar.take(4 * n);
    t.mask = (uint8_t *)J->ar.take(n);
    if (norb >= 2) t.orb = (uint32_t *)J->ar.take(4 * (size_t)norb * norb * n);
    t.kmax = km;
    t.KC = kc;
    t.ccol = (uint16_t *)J->ar.take(16 * n * 2);
    t.cidx = (uint8_t *)J->ar.take(16 * n);
    t.hrec = (uint32_t *)J->ar.take(16 * n * 4);
    *map = pm ? (uint32_t *)J->ar.take(4 * (size_t)norb * n) : nullptr;
    if (J->ar.overflow) return fail(b, EDG_ERR_NOMEM, "problem arena overflow (tables)");
    const double *em_s = J->dparams + (size_t)(spin - 1) * norb;
    const double *ek_s = J->dparams + 2 * norb + (size_t)(spin - 1) * nbath;
    const double *v_s = J->dparams + 2 * norb + 2 * (size_t)nbath + (size_t)(spin - 1) * norb * nbath;
    t.vsrc = v_s;
    // the hop counts are known on the host (km), so no device maximum; the
    // compressed records are written by the same launch
    const double UpmJ = norb >= 2 ? kan_up(J->p) - J->p.J : 0.0;
    if (launch_spin_table(b->side[J->lane], ns, norb, nbath, ne, em_s, ek_s, v_s, J->p.mu, UpmJ, t, true, *map, pm))
        return fail(b, EDG_ERR_HIP, "spin table launch failed");
    return 0;
}

HvOp make_op(const edg_batch *b, const Job *J, int nu, int nd)
{
    HvOp op{};
    const SpinTab &tu = J->tab[0][nu - J->p.nup + 1];
    const SpinTab &td = J->tab[1][nd - J->p.ndn + 1];
    op.nup = tu.n;
    op.ndn = td.n;
    op.norb = J->norb;
    op.nvup = J->norb * J->p.nbath;
    op.ncu = b->ncu;
    op.sched = 0;
    op.up = tu;
    op.dn = td;
    op.U = J->p.U;
    op.Up = J->norb >= 2 ? kan_up(J->p) : 0.0;
    op.Jp = J->norb >= 2 ? kan_jp(J->p) : 0.0;
    return op;
}

// Lay out one chain's vectors and state; the start vector goes in slot 0.
int init_chain(edg_batch *b, Job *J, Chain &C, const HvOp &op, int maxn, bool stored)
{
    const long long D = (long long)op.nup * op.ndn, Dp = rup(D, kPadB);
    BInst &I = C.bi;
    I = BInst{};
    I.op = op;
    I.nslot = stored ? maxn + 1 : 3;
    I.stride = Dp;
    I.base = J->ar.take((size_t)I.nslot * Dp * b->es);
    I.maxn = maxn;
    const Shape sh = chain_shape(b, nsite_of(J->p), op.nup, J->norb);
    I.wide = sh.wide;
    I.cb = sh.cb;
    C.direct = D >= kDirectMin;
    I.hvn = chain_hv_parts(b, nsite_of(J->p), op.nup, op.ndn, J->norb);
    I.axn = chain_ax_parts(D);
    C.ax_n = 0;
    const size_t lds = J->norb >= 2 ? hv2_lds(op.nup, I.cb, b->es) : (size_t)I.cb * op.nup * b->es;
    if (lds > kLdsMax) return fail(b, EDG_ERR_ARG, "tile exceeds LDS");
    char *st = (char *)J->ar.take(state_bytes(maxn, I.hvn, I.axn));
    I.step = (int *)st;
    I.flags = (int *)(st + 8);
    I.alpha = (double *)st + kHead;
    I.beta = I.alpha + maxn;
    I.scale = I.beta + maxn + 1;
    I.hvpart = I.scale + maxn + 1;
    I.axpart = I.hvpart + I.hvn;
    if (J->ar.overflow) return fail(b, EDG_ERR_NOMEM, "problem arena overflow (chain)");
    // zeroed, with flags = {0, maxn}
    if (launch_batch_state_init(b->side[J->lane], st, state_bytes(maxn, I.hvn, I.axn), maxn))
        return fail(b, EDG_ERR_HIP, "state init launch failed");
    C.D = D;
    C.exists = true;
    C.launched = 0;
    return 0;
}


// Returns 1 (nothing changed) when the device cannot hold the problem now.
int admit(edg_batch *b, Job *J)
{
    const edg_problem &p = J->p;
    void *mem = nullptr;
    const int ae = b->heap->alloc(J->bytes, &mem);
    if (ae == 1) return 1;
    if (ae != 0) return fail(b, EDG_ERR_HIP, "device memory allocation failed");
    J->ar.base = (char *)mem;
    J->ar.used = 0;
    J->ar.cap = J->bytes;
    b->in_use += J->bytes;
    J->host = pinned_get(b);
    if (!J->host) return fail(b, EDG_ERR_NOMEM, "pinned host allocation failed");
    for (hipEvent_t *e : {&J->ev, &J->start_ev}) {
        if (!b->ev_free.empty()) {
            *e = b->ev_free.back();
            b->ev_free.pop_back();
        } else {
            BCHK(b, hipEventCreateWithFlags(e, hipEventDisableTiming));
        }
    }
    const int nbath = p.nbath, ns = nsite_of(p), no = J->norb;
    const size_t npar = 2 * (size_t)no + 2 * (size_t)nbath + 2 * (size_t)no * nbath;
    J->dparams = (double *)J->ar.take(8 * npar);
    if (J->ar.overflow) return fail(b, EDG_ERR_NOMEM, "problem arena overflow");
    double *h = J->host;
    if (no >= 2) std::memcpy(h, p.em2, sizeof(double) * 4);
    else {
        h[0] = p.em[0];
        h[1] = p.em[1];
    }
    std::memcpy(h + 2 * no, J->ek.data(), sizeof(double) * 2 * nbath);
    std::memcpy(h + 2 * no + 2 * nbath, J->vmk.data(), sizeof(double) * 2 * no * nbath);
    BCHK(b, hipMemcpyAsync(J->dparams, h, 8 * npar, hipMemcpyHostToDevice, b->side[J->lane]));
    // exact hop counts (zero hybridizations skipped, as the tables do)
    int km[2][3] = {}, kc = 8;
    for (int s = 0; s < 2; ++s) {
        const int n0 = s == 0 ? p.nup : p.ndn;
        for (int d = -1; d <= 1; ++d)
            if (n0 + d >= 0 && n0 + d <= ns) {
                km[s][d + 1] = kmax_exact(no, nbath, n0 + d, J->vmk.data() + (size_t)s * no * nbath);
                if (km[s][d + 1] > 8) kc = 16;
            }
    }
    for (int s = 0; s < 2; ++s) {
        const int n0 = s == 0 ? p.nup : p.ndn;
        for (int d = -1; d <= 1; ++d) {
            const int ne = n0 + d;
            J->has_tab[s][d + 1] = ne >= 0 && ne <= ns;
            if (J->has_tab[s][d + 1])
                if (int rc = build_table(b, J, s + 1, ne, J->tab[s][d + 1], kc, km[s][d + 1],
                                         d == 1 ? 1 : (d == -1 ? 2 : 0), &J->tmap[s][d + 1]))
                    return rc;
        }
    }
    J->misc = (double *)J->ar.take(64 * 8);
    J->coef = (double *)J->ar.take(8 * (8 * (size_t)no * std::max(green_len(p), p.maxnstep) + 64));
    J->gdev = (double *)J->ar.take(8 * 4 * (size_t)no * b->nw);
    const HvOp op = make_op(b, J, p.nup, p.ndn);
    const long long Dp = rup((long long)op.nup * op.ndn, kPadB);
    J->gs = J->ar.take(Dp * b->es);
    if (J->ar.overflow) return fail(b, EDG_ERR_NOMEM, "problem arena overflow");
    if (int rc = init_chain(b, J, J->gsc, op, p.maxnstep, true)) return rc;
    const long long D = J->gsc.D;
    // the start vector with its norm (misc[0]) and scale[0] = 1/|v|
    double *pt = b->side_part[J->lane];
    const int rc0 = b->fp64 ? launch_start<double>(b->side[J->lane], (double *)J->gsc.bi.base, D, p.seed, pt,
                                                   &J->misc[0], J->gsc.bi.scale)
                            : launch_start<float>(b->side[J->lane], (float *)J->gsc.bi.base, D, p.seed, pt,
                                                  &J->misc[0], J->gsc.bi.scale);
    if (rc0) return fail(b, EDG_ERR_HIP, "start launch failed");
    J->gsc.pending = true;  // joins the steps when the host sees start_ev complete
    BCHK(b, hipEventRecord(J->start_ev, b->side[J->lane]));
    J->start_pending = true;
    J->st = ST_GS;
    b->running.push_back(J);
    return 0;
}

// Queue the device->host copy of a finished chain's state block (flags,
// alpha, beta, scale) into the pinned area at off.
int read_chain(edg_batch *b, Job *J, const Chain &C, size_t off)
{
    BCHK(b, hipMemcpyAsync(J->host + off, C.bi.step, sizeof(double) * chain_host_n(C.bi.maxn),
                           hipMemcpyDeviceToHost, b->stream));
    return 0;
}

// a, b of run_lanczos from a state block read back to h (maxn m).
int chain_ab(const double *h, int m, double *a, double *bb)
{
    const int *fl = (const int *)(h + 1);
    const int ns = fl[1];
    const double *ha = h + kHead, *hb = ha + m;
    for (int k = 0; k < m; ++k) {
        a[k] = k < ns ? ha[k] : 0.0;
        bb[k] = (k >= 1 && k <= ns) ? hb[k] : 0.0;
    }
    if (ns < m && ns >= 1) bb[ns] = hb[ns];
    bb[0] = 0.0;
    return ns;
}

// Ground-state chain done: eigenvector, E0 and the c+/c start vectors.
// Host half of the ground-state stage: the lowest eigenvector of the chain's
// tridiagonal matrix, staged as combine coefficients z_j s_j.  Touches only
// the problem's own host data, so several problems run it on threads at
// once.  Returns 0, 1 (breakdown at the start) or 2 (eigensolve failed).
int gs_eigen_body(Job *J)
{
    const int m = J->p.maxnstep;
    std::vector<double> a(m), bb(m);
    const int nst = chain_ab(J->host, m, a.data(), bb.data());
    const double *s = J->host + kHead + 2 * m + 1;
    J->gs_nstep = nst;
    if (nst < 1) return 1;
    std::vector<double> z(nst);
    double lam;
    if (tridiag_lowest(nst, a.data(), bb.data() + 1, &lam, z.data())) return 2;
    double *hc = J->host + chain_host_n(m) + 8;
    for (int j = 0; j < nst; ++j) hc[j] = z[j] * s[j];
    return 0;
}
int gs_eigen(Job *J)
{
    try {
        return gs_eigen_body(J);
    } catch (const std::exception &) {  // host memory: reported as a failed eigensolve
        return 2;
    }
}

// Device half: combine, Rayleigh quotient and the c+/c start vectors.
int after_gs(edg_batch *b, Job *J, int eig_rc)
{
    if (eig_rc == 1) return fail(b, EDG_ERR_SOLVER, "ground-state chain broke down at its start");
    if (eig_rc) return fail(b, EDG_ERR_SOLVER, "tridiagonal eigensolve failed");
    const edg_problem &p = J->p;
    const int m = p.maxnstep, nst = J->gs_nstep;
    double *hc = J->host + chain_host_n(m) + 8;
    BCHK(b, hipMemcpyAsync(J->coef, hc, sizeof(double) * nst, hipMemcpyHostToDevice, b->side[J->lane]));
    const HvOp op = J->gsc.bi.op;
    const long long D = J->gsc.D;
    int rc;
    if (b->fp64) {
        rc = launch_combine<double>(b->side[J->lane], (double *)J->gs, (const double *)J->gsc.bi.base, J->gsc.bi.stride,
                                    nst, J->coef, D, b->side_part[J->lane], &J->misc[4], &J->misc[5]);
        rc |= launch_scale<double>(b->side[J->lane], (double *)J->gs, D, &J->misc[5]);
    } else {
        rc = launch_combine<float>(b->side[J->lane], (float *)J->gs, (const float *)J->gsc.bi.base, J->gsc.bi.stride,
                                   nst, J->coef, D, b->side_part[J->lane], &J->misc[4], &J->misc[5]);
        rc |= launch_scale<float>(b->side[J->lane], (float *)J->gs, D, &J->misc[5]);
    }
    if (rc) return fail(b, EDG_ERR_HIP, "combine launch failed");
    HvArgs ha{};
    ha.x = J->gs;
    ha.partial = b->side_part[J->lane];
    ha.result = &J->misc[1];
    const int hrc = b->fp64 ? launch_hv<double>(b->side[J->lane], op, HV_RQ, 2, true, ha)
                            : launch_hv<float>(b->side[J->lane], op, HV_RQ, 2, true, ha);
    if (hrc < 0) return fail(b, EDG_ERR_HIP, "Rayleigh-quotient launch failed");
    // c+/c start vectors of every orbital (dmft_green.f90:71-118, 202-257),
    // one ring per chain
    const int ns_site = nsite_of(p), mg = green_len(p);
    const bool spin2 = runs_spin2(b, p);
    for (int k = 0; k < kMaxChains; ++k) {
        Chain &C = J->ch[k];
        C.iorb = k / 4 + 1;
        C.ispin = (k & 3) < 2 ? 1 : 2;
        C.pm = (k & 1) ? 2 : 1;
        C.exists = false;
        if (p.gs_only || C.iorb > J->norb) continue;
        if (C.ispin == 2 && !spin2) continue;
        const int nu = p.nup + (C.ispin == 1 ? (C.pm == 1 ? 1 : -1) : 0);
        const int nd = p.ndn + (C.ispin == 2 ? (C.pm == 1 ? 1 : -1) : 0);
        if (nu < 0 || nd < 0 || nu > ns_site || nd > ns_site) continue;
        const HvOp cop = make_op(b, J, nu, nd);
        if (int r = init_chain(b, J, C, cop, mg, false)) return r;
        // the map built with the output sector's table (N + 1 for c+, N - 1
        // for c), at this orbital
        const SpinTab &tout = C.ispin == 1 ? cop.up : cop.dn;
        const uint32_t *map = J->tmap[C.ispin - 1][C.pm == 1 ? 2 : 0] + (size_t)(C.iorb - 1) * tout.n;
        // the start vector c(+) |gs> with |v_init| (misc[8 + k]) and scale[0]
        const int godd = C.ispin == 2 ? (p.nup & 1) : 0;
        double *pt = b->side_part[J->lane];
        const int arc = b->fp64 ? launch_apply_c<double>(b->side[J->lane], (const double *)J->gs, (double *)C.bi.base,
                                                         cop.nup, cop.ndn, op.nup, C.ispin, map, godd, pt,
                                                         &J->misc[8 + k], C.bi.scale)
                                : launch_apply_c<float>(b->side[J->lane], (const float *)J->gs, (float *)C.bi.base,
                                                        cop.nup, cop.ndn, op.nup, C.ispin, map, godd, pt,
                                                        &J->misc[8 + k], C.bi.scale);
        if (arc) return fail(b, EDG_ERR_HIP, "apply_c launch failed");
        C.pending = true;
    }
    // also covers the RQ sums a ground-state-only problem reads next
    BCHK(b, hipEventRecord(J->start_ev, b->side[J->lane]));
    J->start_pending = true;
    J->st = ST_CHAINS;
    return 0;
}

bool chains_running(const Job *J)
{
    for (int k = 0; k < kMaxChains; ++k)
        if (J->ch[k].exists && (J->ch[k].active || J->ch[k].pending)) return true;
    return false;
}

// Chain coefficients -> continued fractions on the device.
int after_chains(edg_batch *b, Job *J)
{
    const edg_problem &p = J->p;
    const int mg = green_len(p), no = J->norb;
    const double *h = J->host;
    const double *rq = h;  // misc[0..15] copied first
    J->e0 = rq[1] / rq[3];
    // per orbital and spin: ap/bp (pm 1) or am/bm (pm 2), b[0] = |v_init|;
    // co[o][s] holds the (mg, 4) block of g_coeffs.dat
    std::vector<double> co[2][2];
    int nsv[2][2][2] = {};
    for (int o = 0; o < no; ++o)
        for (int s = 0; s < 2; ++s) co[o][s].assign(4 * (size_t)mg, 0.0);
    for (int k = 0; k < 4 * no; ++k) {
        const Chain &C = J->ch[k];
        if (!C.exists) continue;
        const double *raw = h + host_chain(k, mg);
        const int o = C.iorb - 1, s = C.ispin - 1, q = C.pm == 1 ? 0 : 2;
        const int ns = chain_ab(raw, mg, co[o][s].data() + q * mg, co[o][s].data() + (q + 1) * mg);
        co[o][s][(q + 1) * mg] = rq[8 + k];
        nsv[o][s][C.pm - 1] = ns;
    }
    const bool spin2 = runs_spin2(b, p);
    if (!spin2)  // the reference reuses the spin-1 coefficients (:102-109)
        for (int o = 0; o < no; ++o) {
            co[o][1] = co[o][0];
            nsv[o][1][0] = nsv[o][0][0];
            nsv[o][1][1] = nsv[o][0][1];
        }
    J->coeffs.resize(8 * (size_t)mg * no);
    for (int o = 0; o < no; ++o)
        for (int s = 0; s < 2; ++s)
            std::memcpy(J->coeffs.data() + ((size_t)o * 2 + s) * 4 * mg, co[o][s].data(), sizeof(double) * 4 * mg);
    double *hc = J->host + host_coef(mg);
    std::memcpy(hc, J->coeffs.data(), sizeof(double) * J->coeffs.size());
    BCHK(b, hipMemcpyAsync(J->coef, hc, sizeof(double) * J->coeffs.size(), hipMemcpyHostToDevice,
                           b->side[J->lane]));
    const int nspin_run = spin2 ? 2 : 1;
    for (int o = 0; o < no; ++o)
        for (int s = 0; s < nspin_run; ++s) {
            const size_t q = ((size_t)o * 2 + s) * 4 * mg;
            const double *dc = J->coef + q;
            if (launch_cf(b->side[J->lane], b->nw, b->omega, J->e0, nsv[o][s][0], dc, dc + mg, nsv[o][s][1], dc + 2 * mg,
                          dc + 3 * mg, J->gdev + ((size_t)o * 2 + s) * 2 * b->nw, 0))
                return fail(b, EDG_ERR_HIP, "continued fraction launch failed");
        }
    double *hg = J->host + host_g(mg);
    BCHK(b, hipMemcpyAsync(hg, J->gdev, sizeof(double) * 4 * b->nw * no, hipMemcpyDeviceToHost, b->side[J->lane]));
    BCHK(b, hipEventRecord(J->ev, b->side[J->lane]));
    J->st = ST_CF_READ;
    return 0;
}

// G_cl(nw, norb, nspin) assembly (dmft_green.f90:111-118) and release.
int finish(edg_batch *b, Job *J)
{
    const edg_problem &p = J->p;
    const int mg = green_len(p), nw = b->nw, no = J->norb;
    const double *hg = J->host + host_g(mg);
    const bool even = !(b->nspin == 2 || p.nup != p.ndn);
    const bool spin2 = runs_spin2(b, p);
    J->G.assign(2 * (size_t)nw * no * b->nspin, 0.0);
    for (int o = 0; o < no; ++o) {
        const double *g1 = hg + (size_t)o * 4 * nw, *g2 = spin2 ? g1 + 2 * nw : g1;
        double *o1 = J->G.data() + 2 * (size_t)o * nw;
        double *o2 = J->G.data() + 2 * ((size_t)no * nw + (size_t)o * nw);
        for (int iw = 0; iw < nw; ++iw) {
            const double r1 = g1[2 * iw], i1 = g1[2 * iw + 1];
            const double r2 = g2[2 * iw], i2 = g2[2 * iw + 1];
            if (b->nspin == 1 && even) {
                o1[2 * iw] = r1;
                o1[2 * iw + 1] = i1;
            } else if (b->nspin == 1) {
                o1[2 * iw] = 0.5 * (r1 + r2);
                o1[2 * iw + 1] = 0.5 * (i1 + i2);
            } else {
                o1[2 * iw] = r1;
                o1[2 * iw + 1] = i1;
                o2[2 * iw] = r2;
                o2[2 * iw + 1] = i2;
            }
        }
    }
    // every GPU user of the region is done: the host saw J->ev after the last
    // stage, and the chains' steps before their coefficients were read
    b->heap->free(J->ar.base);
    J->ar.base = nullptr;
    b->in_use -= J->bytes;
    b->pinned_free.push_back(J->host);
    J->host = nullptr;
    b->ev_free.push_back(J->ev);
    J->ev = nullptr;
    b->ev_free.push_back(J->start_ev);
    J->start_ev = nullptr;
    J->st = ST_DONE;
    b->st_done += 1;
    return 0;
}

// Event-driven stage transitions of every running problem (no blocking).
// Problems whose ground-state chains were read back together (a batch's
// problems finish those chains at the same step) solve their tridiagonal
// matrices on host threads at once, then queue their device work in order.
int progress(edg_batch *b)
{
    std::vector<Job *> gs_ready;
    for (size_t i = 0; i < b->running.size();) {
        Job *J = b->running[i];
        if (J->start_pending) {
            if (hipEventQuery(J->start_ev) != hipSuccess) {
                ++i;
                continue;
            }
            J->start_pending = false;
            Chain *cs[kMaxChains + 1] = {&J->gsc};
            for (int k = 0; k < kMaxChains; ++k) cs[k + 1] = &J->ch[k];
            for (Chain *C : cs)
                if (C->pending) {
                    C->pending = false;
                    C->active = true;
                    if (C->direct) {
                        b->direct.push_back(C);
                        b->moved = true;
                    } else {
                        b->active.push_back(C);
                        b->dirty = true;
                    }
                }
        }
        if (J->st == ST_GS && !J->gsc.active) {
            if (int rc = read_chain(b, J, J->gsc, 0)) return rc;
            BCHK(b, hipEventRecord(J->ev, b->stream));
            J->st = ST_GS_READ;
        } else if (J->st == ST_GS_READ && hipEventQuery(J->ev) == hipSuccess) {
            gs_ready.push_back(J);
        } else if (J->st == ST_CHAINS && !chains_running(J)) {
            // misc (RQ sums, start norms) was written before start_ev
            BCHK(b, hipMemcpyAsync(J->host, J->misc, 16 * sizeof(double), hipMemcpyDeviceToHost, b->stream));
            const int mg = green_len(J->p);
            for (int k = 0; k < kMaxChains; ++k)
                if (J->ch[k].exists)
                    if (int rc = read_chain(b, J, J->ch[k], host_chain(k, mg))) return rc;
            BCHK(b, hipEventRecord(J->ev, b->stream));
            J->st = ST_CH_READ;
        } else if (J->st == ST_CH_READ && hipEventQuery(J->ev) == hipSuccess) {
            if (int rc = after_chains(b, J)) return rc;
        } else if (J->st == ST_CF_READ && hipEventQuery(J->ev) == hipSuccess) {
            if (int rc = finish(b, J)) return rc;
            b->finished.push_back(J);
            b->running.erase(b->running.begin() + i);
            continue;
        }
        ++i;
    }
    const int nr = (int)gs_ready.size();
    if (nr == 0) return 0;
    std::vector<int> erc(nr);
    const int nt = nr < kEigenPerThread ? 1 : std::min(kEigenThreads, (nr + kEigenPerThread - 1) / kEigenPerThread);
    if (nt == 1) {
        for (int k = 0; k < nr; ++k) erc[k] = gs_eigen(gs_ready[k]);
    } else {
        // threads t < started take problems t, t + nt, ...; the host thread
        // takes the rest when the system refuses a thread
        std::vector<std::thread> th;
        int started = 0;
        try {
            for (; started < nt; ++started)
                th.emplace_back([&, t = started]() {
                    for (int k = t; k < nr; k += nt) erc[k] = gs_eigen(gs_ready[k]);
                });
        } catch (const std::exception &) {
        }
        for (int t = started; t < nt; ++t)
            for (int k = t; k < nr; k += nt) erc[k] = gs_eigen(gs_ready[k]);
        for (auto &x : th) x.join();
    }
    for (int k = 0; k < nr; ++k)
        if (int rc = after_gs(b, gs_ready[k], erc[k])) return rc;
    return 0;
}

// Launch class of a chain: single orbital with short columns, long columns
// two per tile, long columns one per tile (x 8 / 16 hop slots); two orbitals
// with 8, 12 or 16 hop slots processed (kan_slots).
int kan_slots(const BInst &I)
{
    const int km = std::max(I.op.up.kmax, I.op.dn.kmax);
    return I.op.up.KC == 8 ? 8 : (km <= 12 ? 12 : 16);
}
int chain_class(const BInst &I)
{
    if (I.op.norb >= 2) return 6 + (kan_slots(I) == 8 ? 0 : (kan_slots(I) == 12 ? 1 : 2));
    return (I.wide ? (I.cb > 1 ? 2 : 4) : 0) + (I.op.up.KC == 16 ? 1 : 0);
}
int class_cols(int c) { return c >= 4 ? 1 : (c >= 2 ? kBatchWideCols : 0); }

// Destroy every graph exec, live and retired.  The caller has drained the
// stream, so no launch of them is still queued.
int drop_graphs(edg_batch *b)
{
    for (hipGraphExec_t &ge : b->gexec) {
        if (!ge) continue;
        if (hipGraphExecDestroy(ge) != hipSuccess) return fail(b, EDG_ERR_HIP, "graph exec destroy failed");
        ge = nullptr;
        b->st_exec_destroyed += 1;
    }
    for (auto &r : b->retired) {
        if (hipGraphExecDestroy(r.first) != hipSuccess) return fail(b, EDG_ERR_HIP, "graph exec destroy failed");
        (void)hipEventDestroy(r.second);
        b->st_exec_destroyed += 1;
    }
    b->retired.clear();
    return 0;
}

// Upload the descriptor table of the active chains grouped by launch class,
// with the tile prefix sums of each class and the axpy prefix sums over all
// chains.  Every chain's tiles start at a multiple of 8 (kXcds): workgroup b
// runs on XCD b % 8, so tile t of a chain runs on XCD t % 8 and the kernels
// deal the chain's column groups to the XCDs in contiguous ranges.  The
// padding tiles exit at once.
int rebuild(edg_batch *b)
{
    std::vector<Chain *> order;
    int start[kClasses + 1] = {};
    for (int c = 0; c < kClasses; ++c) {
        start[c] = (int)order.size();
        for (Chain *C : b->active)
            if (chain_class(C->bi) == c) order.push_back(C);
    }
    start[kClasses] = (int)order.size();
    const int n = start[kClasses];
    std::vector<int> hv[kClasses], ax(n + 1);
    size_t lds[kClasses];  // dynamic LDS bytes of the class's launch
    ax[0] = 0;
    for (int c = 0; c < kClasses; ++c) {
        lds[c] = 16;
        hv[c].assign(start[c + 1] - start[c] + 1, 0);
        for (int i = start[c]; i < start[c + 1]; ++i) {
            const BInst &I = order[i]->bi;
            const int tiles = (I.hvn + 7) / 8 * 8;  // every class: XCD-aligned tile offsets
            hv[c][i - start[c] + 1] = hv[c][i - start[c]] + tiles;
            const size_t l = I.op.norb >= 2 ? hv2_lds(I.op.nup, I.cb, b->es) : (size_t)I.cb * I.op.nup * b->es;
            lds[c] = std::max(lds[c], l);
        }
    }
    for (int i = 0; i < n; ++i) ax[i + 1] = ax[i] + order[i]->bi.axn;
    // chain of every workgroup (a table lookup instead of a search per workgroup)
    std::vector<int> own[kClasses], axown(ax[n]);
    for (int c = 0; c < kClasses; ++c) {
        own[c].resize(hv[c].back());
        for (size_t k = 0; k + 1 < hv[c].size(); ++k)
            for (int w = hv[c][k]; w < hv[c][k + 1]; ++w) own[c][w] = (int)k;
    }
    for (int i = 0; i < n; ++i)
        for (int w = ax[i]; w < ax[i + 1]; ++w) axown[w] = i;
    auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
    size_t off[kClasses], off_own[kClasses];
    size_t total = al(sizeof(BInst) * std::max(n, 1));
    for (int c = 0; c < kClasses; ++c) {
        off[c] = total;
        total += al(hv[c].size() * 4);
        off_own[c] = total;
        total += al(own[c].size() * 4);
    }
    const size_t off_ax = total;
    total += al((size_t)(n + 1) * 4);
    const size_t off_axown = total;
    total += axown.size() * 4;
    const int f = b->flip;
    if (b->hdesc_used[f]) BCHK(b, hipEventSynchronize(b->hdesc_ev[f]));
    if (b->hdesc_cap[f] < total) {
        if (b->hdesc[f]) (void)hipHostFree(b->hdesc[f]);
        BCHK(b, hipHostMalloc((void **)&b->hdesc[f], total * 2, hipHostMallocDefault));
        b->hdesc_cap[f] = total * 2;
    }
    if (b->ddesc_cap < total) {
        // One descriptor buffer for the engine's life, rewritten in stream
        // order; it grows only when the chain set outgrows it.  Captured
        // graphs hold its address, so before it is released the stream
        // drains and every graph exec goes.
        BCHK(b, hipStreamSynchronize(b->stream));
        if (int rc = drop_graphs(b)) return rc;
        if (b->ddesc) BCHK(b, hipFree(b->ddesc));
        b->ddesc = nullptr;
        b->ddesc_cap = 0;
        const size_t cap = std::max(total * 2, kDescMin);
        BCHK(b, hipMalloc((void **)&b->ddesc, cap));
        b->ddesc_cap = cap;
    }
    char *h = b->hdesc[f];
    for (int i = 0; i < n; ++i) std::memcpy(h + sizeof(BInst) * i, &order[i]->bi, sizeof(BInst));
    for (int c = 0; c < kClasses; ++c) {
        std::memcpy(h + off[c], hv[c].data(), 4 * hv[c].size());
        std::memcpy(h + off_own[c], own[c].data(), 4 * own[c].size());
    }
    std::memcpy(h + off_ax, ax.data(), 4 * (size_t)(n + 1));
    std::memcpy(h + off_axown, axown.data(), 4 * axown.size());
    BCHK(b, hipMemcpyAsync(b->ddesc, h, total, hipMemcpyHostToDevice, b->stream));
    BCHK(b, hipEventRecord(b->hdesc_ev[f], b->stream));
    b->hdesc_used[f] = true;
    b->flip ^= 1;
    b->d_inst = (const BInst *)b->ddesc;
    b->d_axoff = (const int *)(b->ddesc + off_ax);
    b->d_axown = (const int *)(b->ddesc + off_axown);
    for (int c = 0; c < kClasses; ++c) {
        b->d_hvoff[c] = (const int *)(b->ddesc + off[c]);
        b->d_hvown[c] = (const int *)(b->ddesc + off_own[c]);
        b->cls_start[c] = start[c];
        b->cls_n[c] = start[c + 1] - start[c];
        b->cls_tiles[c] = hv[c].back();
        b->cls_lds[c] = lds[c];
    }
    b->nax = ax[n];
    b->dirty = false;
    b->st_rebuild += 1;
    b->since_rebuild = 0;
    for (hipGraphExec_t &ge : b->gexec) {
        if (!ge) continue;
        hipEvent_t ev;
        BCHK(b, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        BCHK(b, hipEventRecord(ev, b->stream));
        b->retired.emplace_back(ge, ev);
        ge = nullptr;
    }
    for (size_t i = 0; i < b->retired.size();) {
        const hipError_t q = hipEventQuery(b->retired[i].second);
        if (q == hipSuccess) {
            BCHK(b, hipGraphExecDestroy(b->retired[i].first));
            (void)hipEventDestroy(b->retired[i].second);
            b->retired.erase(b->retired.begin() + i);
            b->st_exec_destroyed += 1;
            continue;
        }
        if (q != hipErrorNotReady) return fail(b, EDG_ERR_HIP, "event query of a retired graph failed");
        ++i;
    }
    return 0;
}

// Queue k batched steps (k <= kChunk): step sub of the chunk is one batched
// H.v launch per class and the axpy, all reading the chains' step + sub;
// one advance after the chunk adds k to every chain's step.
int enqueue_chunk(edg_batch *b, int k)
{
    const int n = (int)b->active.size();
    for (int sub = 0; sub < k; ++sub) {
        for (int c = 0; c < kClasses; ++c) {
            if (!b->cls_n[c]) continue;
            const int rc = c >= 6 ? launch_batch_hv2(b->stream, b->fp64, c == 6 ? 8 : (c == 7 ? 12 : 16),
                                                     b->d_inst + b->cls_start[c],
                                                     b->d_hvoff[c], b->d_hvown[c], b->cls_tiles[c], b->cls_lds[c], sub)
                                  : launch_batch_hv(b->stream, b->fp64, class_cols(c), (c & 1) ? 16 : 8,
                                                    b->d_inst + b->cls_start[c], b->d_hvoff[c], b->d_hvown[c],
                                                    b->cls_tiles[c], (int)(b->cls_lds[c] / b->es), sub);
            if (rc) return fail(b, EDG_ERR_HIP, "batched H.v launch failed");
        }
        if (launch_batch_axpy(b->stream, b->fp64, b->d_inst, b->d_axoff, b->d_axown, b->nax, sub))
            return fail(b, EDG_ERR_HIP, "batched axpy launch failed");
    }
    if (launch_batch_advance(b->stream, b->d_inst, n, k)) return fail(b, EDG_ERR_HIP, "advance launch failed");
    return 0;
}

// Step j of a direct chain: the single-problem H.v and axpy launches of
// edg_cluster_green (lanczos_enqueue in capi.cpp), with that path's default
// schedule.
template <typename T>
int direct_step(edg_batch *b, hipStream_t st, Chain &C, int j)
{
    BInst &I = C.bi;
    T *base = static_cast<T *>(I.base);
    auto slot = [&](int k) { return base + (long long)(k % I.nslot) * I.stride; };
    HvArgs ha{};
    ha.x = slot(j);
    ha.y = slot(j + 1);
    ha.p = j > 0 ? slot(j - 1) : nullptr;
    ha.partial = I.hvpart;
    ha.step = j;
    ha.prev_part = I.axpart;
    ha.prev_n = C.ax_n;
    ha.scale = I.scale;
    ha.beta = I.beta;
    ha.stop = I.flags;
    ha.nstep = I.flags + 1;
    int cols;
    bool acc64;
    default_schedule(I.op.norb, !b->fp64, &cols, &acc64);  // the single-problem path's schedule
    const int hv_n = launch_hv<T>(st, I.op, HV_LANCZOS, cols, acc64, ha);
    if (hv_n < 0 || hv_n > I.hvn) return fail(b, EDG_ERR_HIP, "direct H.v launch failed");
    if (j == I.maxn - 1) {
        if (launch_alpha(st, I.hvpart, hv_n, &I.scale[j], &I.alpha[j], I.flags))
            return fail(b, EDG_ERR_HIP, "direct alpha launch failed");
        return 0;
    }
    const int ax = launch_axpy_norm<T>(st, slot(j + 1), slot(j), C.D, I.hvpart, hv_n, I.scale, I.alpha, j,
                                       I.axpart, I.flags);
    if (ax < 0 || ax > I.axn) return fail(b, EDG_ERR_HIP, "direct axpy launch failed");
    C.ax_n = ax;
    return 0;
}

// k steps of every direct chain, after the batched launches of the chunk
// (outside any graph: their arguments change from chunk to chunk).  With two
// or more, every other chain steps on the last side stream, between a fork
// and a join event, so two chains' launches overlap (each fills the other's
// kernel tails, as the single-problem path's c+ and c chains do).
int run_direct(edg_batch *b, int k)
{
    const bool two = b->direct.size() >= 2;
    hipStream_t s2 = b->side[kSides - 1];
    if (two) {
        BCHK(b, hipEventRecord(b->dfork, b->stream));
        BCHK(b, hipStreamWaitEvent(s2, b->dfork, 0));
    }
    for (size_t i = 0; i < b->direct.size(); ++i) {
        Chain *C = b->direct[i];
        hipStream_t st = two && (i & 1) ? s2 : b->stream;
        for (int sub = 0; sub < k; ++sub) {
            const int rc = b->fp64 ? direct_step<double>(b, st, *C, C->launched + sub)
                                   : direct_step<float>(b, st, *C, C->launched + sub);
            if (rc) return rc;
        }
    }
    if (two) {
        BCHK(b, hipEventRecord(b->djoin, s2));
        BCHK(b, hipStreamWaitEvent(b->stream, b->djoin, 0));
    }
    b->st_steps += k;
    b->st_chain_steps += (double)k * b->direct.size();
    for (size_t i = 0; i < b->direct.size();) {
        Chain *C = b->direct[i];
        C->launched += k;
        if (C->launched >= C->bi.maxn) {
            C->active = false;
            b->direct.erase(b->direct.begin() + i);
            b->moved = true;
            continue;
        }
        ++i;
    }
    return 0;
}

// Run k steps of the active chains (1 <= k <= kChunk, k <= every
// chain's remainder) as one chunk: a captured graph of k steps, captured on
// the first chunk of that length after the set of chains last changed.
// The direct chains follow the chunk with their own launches.
int run_steps(edg_batch *b, int k)
{
    if (k < 1 || k > kChunk) return fail(b, EDG_ERR_ARG, "chunk length out of range");
    if (b->active.empty()) return run_direct(b, k);
    if (b->dirty)
        if (int rc = rebuild(b)) return rc;
    // A graph pays only when it is replayed: chunks right after a change of
    // the chain set (chains start and finish all the time in a served
    // stream) launch their kernels directly, and a chunk length is captured
    // once the set has stayed unchanged for kGraphAfter chunks.
    const bool graph = b->use_graphs && (b->gexec[k] || b->since_rebuild >= kGraphAfter);
    b->since_rebuild += 1;
    if (graph) {
        if (!b->gexec[k]) {
            hipGraph_t g = nullptr;
            BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue_chunk(b, k);
            hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (rc) return rc;
            if (e != hipSuccess) return fail(b, EDG_ERR_HIP, "graph capture failed");
            e = hipGraphInstantiate(&b->gexec[k], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                b->gexec[k] = nullptr;
                return fail(b, EDG_ERR_HIP, "graph instantiation failed");
            }
            b->st_inst += 1;
        }
        BCHK(b, hipGraphLaunch(b->gexec[k], b->stream));
        b->st_graph += 1;
    } else if (int rc = enqueue_chunk(b, k)) {
        return rc;
    }
    if (int rc = run_direct(b, k)) return rc;
    b->st_chain_steps += (double)k * b->active.size();
    for (size_t i = 0; i < b->active.size();) {
        Chain *C = b->active[i];
        C->launched += k;
        if (C->launched >= C->bi.maxn) {
            C->active = false;
            b->active.erase(b->active.begin() + i);
            b->dirty = true;
            continue;
        }
        ++i;
    }
    return 0;
}

// Admission in submission order by device bytes: a large problem waits
// while kMaxLarge others run, and the problems behind it go ahead when they
// fit.  The device memory of the queue's problems that fit the budget is
// reserved first, in one piece: growing the heap chunk by chunk while the
// device runs cost up to a second per batch.
int admit_queued(edg_batch *b)
{
    const auto t = std::chrono::steady_clock::now();
    size_t want = 0;
    for (const Job *J : b->queue) {
        if (b->in_use + want + J->bytes > b->budget) break;
        want += J->bytes;
    }
    if (want > 0 && b->heap->reserve(want) < 0) return fail(b, EDG_ERR_HIP, "device memory reservation failed");
    int large_in_flight = 0;
    for (const Job *R : b->running) large_in_flight += R->large ? 1 : 0;
    for (auto it = b->queue.begin(); it != b->queue.end();) {
        Job *J = *it;
        if (J->large && large_in_flight >= kMaxLarge) {
            ++it;
            continue;
        }
        if (b->in_use + J->bytes > b->budget) break;
        const int rc = admit(b, J);
        if (rc == 1) {  // device full: wait for running problems to release memory
            if (b->running.empty())
                return fail(b, EDG_ERR_NOMEM, "device memory exhausted with no problem in flight");
            break;
        }
        if (rc) return rc;
        large_in_flight += J->large ? 1 : 0;
        it = b->queue.erase(it);
    }
    b->st_t_admit += std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
    return 0;
}

}  // namespace

extern "C" {

int edg_batch_create(int device, int precision, int nspin, int nw, const double *omega,
                     unsigned long long budget_bytes, edg_batch **out)
{
    if (!out) return EDG_ERR_ARG;
    *out = nullptr;
    if ((precision != EDG_FP32 && precision != EDG_FP64) || (nspin != 1 && nspin != 2) || nw < 1 || !omega)
        return EDG_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return EDG_ERR_HIP;
    edg_batch *b = new edg_batch;
    b->device = device;
    b->fp64 = precision == EDG_FP64;
    b->es = b->fp64 ? 8 : 4;
    b->nspin = nspin;
    b->nw = nw;
    b->pinned_n = host_g(kMaxLen) + 8 * (size_t)nw;
    int rc = 0;
    auto chk = [&](hipError_t e) {
        if (e != hipSuccess && !rc) rc = EDG_ERR_HIP;
    };
    chk(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    for (hipStream_t &s : b->side) chk(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    for (auto &e : b->chunk_ev) chk(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    chk(hipEventCreateWithFlags(&b->dfork, hipEventDisableTiming));
    chk(hipEventCreateWithFlags(&b->djoin, hipEventDisableTiming));
    chk(hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, device));
    chk(hipMalloc((void **)&b->omega, sizeof(double) * nw));
    for (double *&p : b->side_part) chk(hipMalloc((void **)&p, sizeof(double) * kPartialCap));
    chk(hipEventCreateWithFlags(&b->hdesc_ev[0], hipEventDisableTiming));
    chk(hipEventCreateWithFlags(&b->hdesc_ev[1], hipEventDisableTiming));
    if (!rc) chk(hipMemcpy(b->omega, omega, sizeof(double) * nw, hipMemcpyHostToDevice));
    b->heap = new DevHeap(device);
    if (!rc && batch_set_lds(kLdsMax)) rc = EDG_ERR_HIP;
    if (!rc && budget_bytes == 0) {
        // free device memory plus the idle chunks earlier queues left in the
        // device cache (hipMemGetInfo counts those as used)
        size_t fr = 0, tot = 0;
        chk(hipMemGetInfo(&fr, &tot));
        budget_bytes = (unsigned long long)(0.85 * (double)(fr + devmem_cached_bytes(device)));
    }
    if (rc) {
        (void)hipGetLastError();
        edg_batch_destroy(b);
        return rc;
    }
    b->budget = budget_bytes;
    *out = b;
    return EDG_OK;
}

int edg_batch_destroy(edg_batch *b)
{
    if (!b) return EDG_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (hipStream_t s : b->side)
        if (s) (void)hipStreamSynchronize(s);
    auto drop = [&](Job *J) {
        if (J->ar.base && b->heap) b->heap->free(J->ar.base);
        if (J->ev) (void)hipEventDestroy(J->ev);
        if (J->start_ev) (void)hipEventDestroy(J->start_ev);
        delete J;
    };
    for (Job *J : b->queue) drop(J);
    for (Job *J : b->running) drop(J);
    for (Job *J : b->finished) drop(J);
    for (const PinSlab &sl : b->slabs) {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        if (g_pin_cached + sl.bytes <= kPinCacheMax) {
            g_pin_cache.push_back(sl);
            g_pin_cached += sl.bytes;
        } else {
            (void)hipHostFree(sl.p);
        }
    }
    for (hipEvent_t e : b->ev_free) (void)hipEventDestroy(e);
    (void)drop_graphs(b);
    if (b->ddesc) (void)hipFree(b->ddesc);
    delete b->heap;  // its chunks go to the device cache for later queues
    for (int f = 0; f < 2; ++f) {
        if (b->hdesc[f]) (void)hipHostFree(b->hdesc[f]);
        if (b->hdesc_ev[f]) (void)hipEventDestroy(b->hdesc_ev[f]);
    }
    if (b->omega) (void)hipFree(b->omega);
    for (double *p : b->side_part)
        if (p) (void)hipFree(p);
    for (auto &e : b->chunk_ev)
        if (e) (void)hipEventDestroy(e);
    if (b->dfork) (void)hipEventDestroy(b->dfork);
    if (b->djoin) (void)hipEventDestroy(b->djoin);
    for (hipStream_t s : b->side)
        if (s) (void)hipStreamDestroy(s);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return EDG_OK;
}

int edg_batch_release_memory(int device)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return EDG_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        for (size_t i = 0; i < g_pin_cache.size();)
            if (g_pin_cache[i].device == device) {
                (void)hipHostFree(g_pin_cache[i].p);
                g_pin_cached -= g_pin_cache[i].bytes;
                g_pin_cache.erase(g_pin_cache.begin() + i);
            } else {
                ++i;
            }
    }
    return devmem_release_cache(device) ? EDG_ERR_HIP : EDG_OK;
}

const char *edg_batch_last_error(const edg_batch *b) { return b ? b->err.c_str() : "null batch"; }

int edg_batch_set_graphs(edg_batch *b, int use_graphs)
{
    if (!b) return EDG_ERR_ARG;
    b->use_graphs = use_graphs != 0;
    return EDG_OK;
}

int edg_batch_set_schedule(edg_batch *b, int n, const int *nsite_max, const int *tile_elems, const int *wide)
{
    if (!b || n < 1 || n > 16 || !nsite_max || !tile_elems || !wide) return EDG_ERR_ARG;
    if (!b->queue.empty() || !b->running.empty())
        return fail(b, EDG_ERR_STATE, "the schedule can change only while no problem is queued or in flight");
    std::vector<edg_batch::Sched> s;
    for (int i = 0; i < n; ++i) {
        if (tile_elems[i] < 64 || tile_elems[i] > 16384 || (wide[i] != 0 && wide[i] != 1) || nsite_max[i] < 1 ||
            (i > 0 && nsite_max[i] <= nsite_max[i - 1]))
            return fail(b, EDG_ERR_ARG, "schedule entries: increasing nsite_max, tile 64..16384, wide 0/1");
        s.push_back({nsite_max[i], tile_elems[i], wide[i]});
    }
    b->sched = s;
    return EDG_OK;
}

int edg_batch_problem_bytes(const edg_batch *b, const edg_problem *p, unsigned long long *bytes)
{
    if (!b || !p || !bytes) return EDG_ERR_ARG;
    if (int rc = validate(b, *p)) return rc;
    *bytes = problem_bytes(b, *p);
    return EDG_OK;
}

int edg_batch_submit(edg_batch *b, const edg_problem *p, long long *id)
{
    if (!b || !p || !id) return EDG_ERR_ARG;
    if (int rc = validate(b, *p)) return rc;
    Job *J = new Job;
    J->p = *p;
    J->norb = norb_of(*p);
    J->ek.assign(p->ek, p->ek + 2 * (size_t)p->nbath);
    J->vmk.assign(p->vmk, p->vmk + 2 * (size_t)J->norb * p->nbath);
    J->p.ek = nullptr;
    J->p.vmk = nullptr;
    J->bytes = problem_bytes(b, *p);
    {
        const int ns = nsite_of(*p);
        J->large = (size_t)(binom_ll(ns, p->nup) * binom_ll(ns, p->ndn)) * b->es >= kLargeVec;
    }
    if (J->bytes > b->budget) {
        delete J;
        return fail(b, EDG_ERR_ADMISSION, "problem needs " + std::to_string(problem_bytes(b, *p)) +
                                              " device bytes, budget is " + std::to_string(b->budget));
    }
    J->id = b->next_id++;
    J->lane = b->next_lane;
    b->next_lane = (b->next_lane + 1) % kSides;
    *id = J->id;
    b->queue.push_back(J);
    return EDG_OK;
}

int edg_batch_step(edg_batch *b, int max_steps, int *pending)
{
    if (!b) return EDG_ERR_ARG;
    (void)hipSetDevice(b->device);
    using clk = std::chrono::steady_clock;
    auto secs = [](clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); };
    auto t = clk::now();
    if (int rc = progress(b)) return rc;
    b->st_t_progress += secs(t);
    if (int rc = admit_queued(b)) return rc;
    if ((!b->active.empty() || !b->direct.empty()) && max_steps > 0) {
        int left = max_steps;
        while (left > 0 && (!b->active.empty() || !b->direct.empty())) {
            int k = std::min(left, kChunk);
            for (const Chain *C : b->active) k = std::min(k, C->bi.maxn - C->launched);
            for (const Chain *C : b->direct) k = std::min(k, C->bi.maxn - C->launched);
            b->moved = false;
            t = clk::now();
            if (int rc = run_steps(b, k)) return rc;
            left -= k;
            // bounded run-ahead: wait for the chunk kAhead chunks back
            BCHK(b, hipEventRecord(b->chunk_ev[b->chunk_head], b->stream));
            b->chunk_head = (b->chunk_head + 1) % kChunkEvents;
            b->st_t_steps += secs(t);
            if (++b->chunk_n > kAhead) {
                const int old = (b->chunk_head + kChunkEvents - b->chunk_n) % kChunkEvents;
                t = clk::now();
                BCHK(b, hipEventSynchronize(b->chunk_ev[old]));
                b->st_t_wait += secs(t);
                --b->chunk_n;
            }
            t = clk::now();
            if (int rc = progress(b)) return rc;
            b->st_t_progress += secs(t);
            if (b->dirty || b->moved) break;  // chains joined or left: let the caller see progress
        }
    } else if (!b->running.empty()) {
        // nothing to step: wait for the oldest outstanding event, then move on
        t = clk::now();
        for (Job *J : b->running) {
            if (J->start_pending) {
                BCHK(b, hipEventSynchronize(J->start_ev));
                break;
            }
            if (J->st == ST_GS_READ || J->st == ST_CH_READ || J->st == ST_CF_READ) {
                BCHK(b, hipEventSynchronize(J->ev));
                break;
            }
        }
        b->st_t_wait += secs(t);
        t = clk::now();
        if (int rc = progress(b)) return rc;
        b->st_t_progress += secs(t);
    }
    if (pending) *pending = (int)(b->queue.size() + b->running.size());
    return EDG_OK;
}

int edg_batch_run(edg_batch *b)
{
    if (!b) return EDG_ERR_ARG;
    int pending = 1;
    while (pending > 0)
        if (int rc = edg_batch_step(b, 1 << 20, &pending)) return rc;
    return EDG_OK;
}

int edg_batch_poll(edg_batch *b, long long *id, double *e0, double *G, int *nstep, double *coeffs)
{
    if (!b) return EDG_ERR_ARG;
    if (b->finished.empty()) return 0;
    Job *J = b->finished.front();
    b->finished.pop_front();
    if (id) *id = J->id;
    if (e0) *e0 = J->e0;
    if (nstep) *nstep = J->gs_nstep;
    if (G) std::memcpy(G, J->G.data(), sizeof(double) * J->G.size());
    if (coeffs) std::memcpy(coeffs, J->coeffs.data(), sizeof(double) * J->coeffs.size());
    delete J;
    return 1;
}

int edg_batch_stats(const edg_batch *b, double *stats)
{
    if (!b || !stats) return EDG_ERR_ARG;
    stats[0] = b->st_steps;
    stats[1] = b->st_chain_steps;
    stats[2] = b->st_done;
    stats[3] = (double)b->in_use;
    stats[4] = (double)b->budget;
    stats[5] = (double)b->running.size();
    stats[6] = b->st_graph;
    stats[7] = b->st_rebuild;
    stats[8] = b->st_t_admit;
    stats[9] = b->st_t_progress;
    stats[10] = b->st_t_steps;
    stats[11] = b->st_t_wait;
    return EDG_OK;
}

int edg_batch_graph_stats(const edg_batch *b, long long *instantiated, long long *destroyed, int *live)
{
    if (!b) return EDG_ERR_ARG;
    int n = (int)b->retired.size();
    for (hipGraphExec_t ge : b->gexec) n += ge != nullptr;
    if (instantiated) *instantiated = b->st_inst;
    if (destroyed) *destroyed = b->st_exec_destroyed;
    if (live) *live = n;
    return EDG_OK;
}

}  // extern "C"
