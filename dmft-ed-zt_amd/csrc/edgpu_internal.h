// Internal declarations shared by the HIP kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace edg {

constexpr int kBlock = 256;       // threads per workgroup of the vector kernels
constexpr int kHvBlock = 512;     // threads per workgroup of the H.v kernel (8 wave64)
constexpr int kMaxHops = 64;      // per-spin hop slots per config (norb*nbath)
constexpr int kMinSlots = 16;     // tables hold at least this many slots (padding)
constexpr int kMaxNorb = 5;       // observables accumulate 3*norb sums
constexpr uint32_t kInvalid = 0xFFFFFFFFu;

// One hop of a per-spin table: target rank and matrix element.  Unused
// slots point at the row itself with a zero element, so that unrolled
// loops can read every slot without bounds tests.
struct HopF {
    uint32_t col;
    float val;
};
struct HopD {
    uint32_t col;
    uint32_t pad;
    double val;
};

// Device arrays describing one spin half (spin sigma, N_sigma electrons) of
// a sector under one set of Hamiltonian parameters.
struct SpinTab {
    int n = 0;                     // C(nsite, N)
    int K = 0;                     // slots per config = max(norb*nbath, kMinSlots)
    int kmax = 0;                  // largest hop count over all configs
    uint32_t *pattern = nullptr;   // [n]
    HopF *hf = nullptr;            // [K][n] (FP32 element)
    HopD *hd = nullptr;            // [K][n] (FP64 element)
    uint8_t *cnt = nullptr;        // [n]
    double *energy = nullptr;      // [n]
    float *energyf = nullptr;      // [n] FP32 copy for FP32 row sums
    uint8_t *mask = nullptr;       // [n] occupied impurity orbitals
    uint32_t *orb = nullptr;       // [norb*norb][n] orbital moves dst<-src
    int *maxcnt = nullptr;         // device scalar: max over cnt (null: not computed)
    // compressed hop tables of the single-orbital kernels, [n][KC]:
    // ccol = target rank, cidx = index into vsrc (bit 7: odd Jordan-Wigner
    // parity, the element is -vsrc[idx]), 64 = zero; hrec = target rank |
    // (value-table index << 24), bits 16-23 zero (rec_rank / rec_vidx in
    // device_util.h), the value table being [idx] = +v,
    // [kMaxHops + 1 + idx] = -v (v the signed hybridization), idx 64 = 0
    int KC = 0;                    // 8 or 16 (0: not built)
    uint16_t *ccol = nullptr;
    uint8_t *cidx = nullptr;
    uint32_t *hrec = nullptr;
    const double *vsrc = nullptr;  // hybridizations of this spin, [l*norb + m]
};

// Coefficients of fused Lanczos step j (HV_FUSED), in FP64 and FP32:
//   [0] s_j  [1] d = alpha_{j-1} - sigma_{j-1}  [2] d s_{j-1}
//   [3] sigma_{j-1} s_{j-1}  [4] beta_{j-1} s_{j-2}  [5] sigma_j s_j
//   [6] beta_j s_{j-1}
// written by the finisher of step j-1 (by the start-vector norm for j = 0)
// so that the kernel reads them straight into scalar registers.
struct FusedCoef {
    double d[8];
    float f[8];
};

// Reorthogonalization of one Lanczos chain after its axpy (reorth.hip).
struct ReorthArgs {
    int mode;              // 1 = partial (omega recurrence), 2 = full (every step)
    int step;              // j: r = unnormalized u_{j+1}, basis slots 0..j
    int ld;                // row length of omega / dpart / coef (>= maxn + 2)
    int ax_n;              // partials of |r|^2 the next H.v reduces (the axpy's count)
    double eps;            // unit roundoff of the stored vectors
    const double *alpha, *beta, *scale;
    double *ax_part;       // |r|^2 partials: read, and rewritten when the step reorthogonalizes
    double *omega;         // 3 rows of ld: omega rows j-1, j, j+1 (ring by j % 3)
    int *flags;            // [0] this step reorthogonalizes, [2] count, [3 + j % 2] step j is forced,
                           // [5] steps run
    const int *stop;       // the chain's stop flag
    double *dpart;         // [vector][workgroup] partial dots
    double *coef;          // [ld] c_k = s_k^2 <u_k, r>
};
// Workgroups of the partial dots; the buffer holds kReorthDots * ld doubles.
constexpr int kReorthDots = 2048;
template <typename T>
int launch_reorth(hipStream_t s, const ReorthArgs &a, T *r, const T *base, long long stride, long long n,
                  int ndot);

__host__ __device__ constexpr size_t align16(size_t n) { return (n + 15) & ~size_t(15); }

constexpr int kNorb2 = 2;
constexpr int kPairs2 = kNorb2 * (kNorb2 - 1);

// Dynamic LDS of the two-orbital H.v tile: the C columns and their C target
// columns.
__host__ __device__ constexpr size_t hv2_lds(int nup, int C, size_t esize)
{
    return 2 * align16((size_t)nup * C * esize);
}

// Everything the H.v kernel needs besides the vectors.
struct HvOp {
    int nup, ndn, norb;
    int nvup;  // hybridization elements of the up spin (norb*nbath)
    int ncu;   // compute units of the device (persistent grid size)
    int sched; // 0: one workgroup per column group; 16/32/64: persistent, C = sched/16
    SpinTab up, dn;
    double U, Up, Jp;
};

// HV_PLAIN: y = Hx (y += Hx when HvArgs::accumulate; the launcher then runs
// the HV_ACCUM instantiation, so that the plain form never reads y).
// HV_FUSED: one Lanczos step without a separate axpy pass (single-orbital
// tile only).  Step j gathers from the shifted residual of step j-1,
// r_{j-1} = (H - sigma_{j-1}) q_{j-1} - beta_{j-1} q_{j-2} (u_0 at j = 0),
// where q_k = s_k u_k and u_k are the stored unnormalized Krylov vectors,
// and per element forms (d = alpha_{j-1} - sigma_{j-1})
//   u_j   = r_{j-1} - d q_{j-1}                                        (stored)
//   H u_j = H r_{j-1} - d (r_{j-1} + sigma_{j-1} q_{j-1} + beta_{j-1} q_{j-2})
//   r_j   = s_j H u_j - sigma_j q_j - beta_j q_{j-1}                   (stored)
// with partial sums <u_j, r_j>, |r_j|^2, |u_j|^2, from which the finisher
// takes alpha_j and beta_{j+1}.  H u_j is never applied to the stored u_j:
// it is carried by the identity in the middle line, so an error in it is
// multiplied by d / beta_j at every step.  Without the shift (sigma = 0)
// that factor is alpha / beta, large for chains that start low in the
// spectrum (measured FP64 errors of 5e-8 in G); with sigma_j = alpha_{j-1}
// it is the step-to-step change of alpha over beta.
enum HvMode { HV_PLAIN = 0, HV_LANCZOS = 1, HV_RQ = 2, HV_ACCUM = 3, HV_FUSED = 4 };

struct HvArgs {
    const void *x;          // input vector (T)
    void *y;                // output vector (T)
    const void *p;          // previous Lanczos vector (T) or null
    int accumulate;         // HV_PLAIN: y += Hx
    double *partial;        // per-block partial sums written by this kernel
    double *result;         // HV_RQ: the three sums
    // HV_LANCZOS step j: s_j = 1/beta_j with beta_j^2 = sum of prev_part
    // (the axpy partials of step j-1; step 0 uses scale[0]); workgroup 0
    // records beta[j] and scale[j]; beta_j < 1e-16 stops the chain
    int step;
    const double *prev_part;
    int prev_n;
    double *scale;
    double *beta;
    int *stop;              // skip when *stop != 0
    int *nstep;
    // HV_FUSED: x = r'_{j-1} (u_0 at j = 0), p = u_{j-1}, p2 = u_{j-2},
    // u = where u_j goes (null at j = 0), y = r'_j; alpha, beta, scale the
    // chain's coefficient arrays (alpha_{j-1}, beta_{j-1}, beta_j, s_{j-2..j}
    // written by the finishers of the earlier steps)
    const void *p2;
    void *u;
    double *alpha;
    struct FusedCoef *fcoef;  // [step]: the step's coefficients, written by the previous finisher
};

// One Lanczos chain of the batched engine (batch.hip / batch.cpp): its
// sector operator, its vectors and its device-resident step state.  Vector
// slot k is base + (k % nslot) * stride (nslot = maxn + 1 for a stored
// Krylov basis, 3 for a ring).  The step index lives on the device and is
// advanced by a kernel, so a batched step is the same three launches every
// time (capturable in a HIP graph).
struct BInst {
    HvOp op;
    void *base;
    long long stride;
    int nslot;
    int maxn;
    int wide;          // 1: one long column per 512-thread tile; 0: cb columns per 256-thread tile
    int cb;            // sector columns per H.v tile
    int hvn, axn;      // H.v tiles / axpy workgroups of this chain
    int *step;         // step of the current chunk's first step (a launch adds its offset)
    int *flags;        // [0] stop, [1] nstep
    double *alpha;     // [maxn]
    double *beta;      // [maxn + 1]
    double *scale;     // [maxn + 1]
    double *hvpart;    // [hvn]
    double *axpart;    // [axn]
};

// Largest workgroup count of launch_axpy_norm (its partial sums).
constexpr int kAxpyMaxGrid = 256;

// ---- launchers (kernels.hip) ----
int launch_spin_table(hipStream_t s, int nsite, int norb, int nbath, int ne, const double *em_s,
                      const double *ek_s, const double *v_s, double mu, double UpmJ,
                      const SpinTab &t, bool compress = false, uint32_t *cmap = nullptr,
                      int cmap_pm = 0);
int launch_compress(hipStream_t s, const SpinTab &t);
// Returns the number of workgroups (= partial sums written) or < 0 on error.
template <typename T>
int launch_hv(hipStream_t s, const HvOp &op, int mode, int cols, bool acc64, const HvArgs &a);
// Default H.v schedule of a new context (and of edg_hx_noalloc), measured on
// MI355X at Ns=14, sector (7,7): single orbital, two columns per workgroup
// (FP32 43 us, FP64 81 us) beat one (51 us, 87 us); two orbitals, FP32
// vectors: two columns with FP32 row sums (H.v 65.6 us, solve 122 ms) beat
// FP64 sums (75.1 us, 131 ms) and one column (73.2 us, 130 ms), with errors
// against the FP64 oracle's solve of |dE0| 1.7e-13 and max|dG| 2.5e-7 (FP64
// sums: 1.3e-11 and 7.9e-7); FP64 vectors: one column (solve 262 against
// 280 ms).  The general kernel (norb >= 3) sums FP32 rows in FP64.
inline void default_schedule(int norb, bool fp32, int *cols, bool *acc64)
{
    *cols = (norb == 1 || (norb == 2 && fp32)) ? 2 : 1;
    *acc64 = fp32 && norb >= 3;
}
// alpha[step] = scale[step] * sum(partial[0..nblk)) unless stopped.
int launch_alpha(hipStream_t s, const double *partial, int nblk, const double *scale_j,
                 double *alpha_j, int *stop);
// Finisher of an HV_FUSED step: alpha[step], beta[step + 1], scale[step + 1]
// from the three sums per workgroup; beta < 1e-16 stops the chain
// (*nstep = step + 1), as FIN_BETA does.
int launch_fused_finish(hipStream_t s, const double *partial, int nblk, int step, double *alpha,
                        double *beta, double *scale, int *stop, int *nstep, FusedCoef *fc);
// fc[0] of a fused chain from scale[0] (after the start-vector norm).
int launch_fused_start(hipStream_t s, const double *scale, FusedCoef *fc);
// Whether launch_hv runs HV_FUSED for this operator and schedule.
bool hv_fused_supported(const HvOp &op, int cols, size_t esize);
template <typename T>
int launch_norm(hipStream_t s, const T *x, long long n, double *partial,
                double *norm_out, double *scale_out, const int *stop);
// r <- r - alpha_j s_j u_j with alpha_j = s_j * sum(hv_part) (recorded in
// alpha[step]); writes partial sums of |r|^2.  Returns the workgroup count.
template <typename T>
int launch_axpy_norm(hipStream_t s, T *r, const T *u, long long n, const double *hv_part,
                     int hv_n, const double *scale, double *alpha, int step, double *partial,
                     int *stop);
// partial non-null: also |x| into norm_out and 1/|x| into scale_out (the
// sums of launch_norm, in its order).
template <typename T>
int launch_start(hipStream_t s, T *x, long long n, unsigned long long seed, double *partial = nullptr,
                 double *norm_out = nullptr, double *scale_out = nullptr);
// Streaming reference passes over n floats (n a multiple of 4): y = x
// (copy) and y = x + z (add), 16-byte accesses, grid-stride.
int launch_stream_copy(hipStream_t s, const float *x, float *y, long long n);
int launch_stream_add(hipStream_t s, const float *x, const float *z, float *y, long long n);
template <typename T>
int launch_combine(hipStream_t s, T *out, const T *basis, long long stride, int nvec,
                   const double *coef, long long n, double *partial,
                   double *norm_out, double *scale_out);
template <typename T>
int launch_scale(hipStream_t s, T *x, long long n, const double *scale);
template <typename T>
int launch_from_f64(hipStream_t s, T *dst, const double *src, long long n);
template <typename T>
int launch_to_f64(hipStream_t s, double *dst, const T *src, long long n);
int launch_cmap(hipStream_t s, const uint32_t *pattern_out, int n_out, int nsite, int ne_in,
                int bit, int pm, uint32_t *map);
// partial non-null: also |out| and 1/|out|, as launch_start.
template <typename T>
int launch_apply_c(hipStream_t s, const T *in, T *out, int nup_out, int ndn_out, int nup_in,
                   int spin, const uint32_t *map, int global_odd, double *partial = nullptr,
                   double *norm_out = nullptr, double *scale_out = nullptr);
template <typename T>
int launch_observables(hipStream_t s, const T *gs, const HvOp &op, const double *scale,
                       double *partial, double *out);
int launch_sigma(hipStream_t s, int nw, int norb, int nbath, int nspin, const double *omega,
                 const double *params, double mu, const double *G, double *Sigma, double *Z);
int launch_cf(hipStream_t s, int nw, const double *omega, double e0, int np, const double *ap,
              const double *bp, int nm, const double *am, const double *bm, double *g,
              int accumulate);

// ---- batched Lanczos steps (batch.hip) ----
// hv_off / ax_off: prefix sums (ninst + 1 entries) of the chains' H.v tiles
// and axpy workgroups; the launches cover hv_off[ninst] / ax_off[ninst]
// workgroups, and own[w] is the chain of workgroup w.  lds_elems: largest
// staged elements per tile over the chains.
// wide_cols: 0 for short-column chains (BInst::cb columns per 256-thread
// tile); 1 or kBatchWideCols for long-column chains (that many columns per
// 512-thread tile, tile offsets of every chain a multiple of 8).  ks: hop
// slots per row of every table of the launched chains (8 or 16).
// The batched kernels run step *step + sub of every chain; the advance adds
// k to every chain's step after a chunk of k steps (sub = 0 .. k-1).
int launch_batch_hv(hipStream_t s, bool fp64, int wide_cols, int ks, const BInst *inst, const int *hv_off,
                    const int *own, int ntiles, int lds_elems, int sub);
// Two-orbital chains: kBatchKanColsF32 (FP32 vectors, FP64 accumulation) or
// 1 (FP64) columns per 512-thread tile, lds_bytes = hv2_lds of the largest
// nup among them; kh = hop slots processed per row (8 with 8-slot tables,
// 12 or 16 with 16-slot tables), at least every table's largest hop count.
int launch_batch_hv2(hipStream_t s, bool fp64, int kh, const BInst *inst, const int *hv_off, const int *own,
                     int ntiles, size_t lds_bytes, int sub);
constexpr int kBatchKanColsF32 = 2;
int launch_batch_axpy(hipStream_t s, bool fp64, const BInst *inst, const int *ax_off, const int *own,
                      int nblocks, int sub);
int launch_batch_advance(hipStream_t s, const BInst *inst, int ninst, int k);
int launch_batch_state_init(hipStream_t s, void *st, size_t bytes, int maxn);
// Allow max_lds bytes of dynamic LDS for the batched H.v (call once, outside
// any graph capture).
int batch_set_lds(size_t max_lds);
// Elements per batched H.v tile: cb = clamp(kBatchTile / nup, 1, kBatchMaxCols)
// columns; a tile stages cb * nup elements in LDS.
constexpr int kBatchTile = 2048;
constexpr int kBatchMaxCols = 64;
constexpr int kBatchWide = 1024;
constexpr int kBatchWideCols = 2;  // columns per long-column tile when they fit LDS

// Number of per-block partial slots the reductions may need.

}  // namespace edg
