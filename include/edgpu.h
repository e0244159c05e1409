/*
 * edgpu.h -- C ABI of the MI355X exact-diagonalization impurity solver.
 *
 * Plain extern "C" entry points with pointers and sizes only.  A Fortran
 * caller binds them through iso_c_binding (dmft-ed-zt_amd/fortran/edgpu_c.f90
 * declares every entry point with bind(C)); dmft-ed-zt_amd/fortran/
 * edgpu_dropin.f90 re-exposes them with the reference's argument lists.  Each entry point names the reference routine it replaces
 * (paths relative to the yw-choi/DMFT-ED-ZT tree).
 *
 * Array conventions (identical to the reference):
 *   em(norb,2), ek(nbath,2), vmk(norb,nbath,2): FP64, column-major;
 *   sector vectors: FP64, length C(nsite,N_up)*C(nsite,N_dn), row index
 *   (rank_dn-1)*nup + rank_up (up rank fastest), ranks in increasing
 *   bit-pattern order (ed_basis.f90:42-151);
 *   complex arrays: interleaved (re, im) FP64, Fortran double complex layout.
 *
 * Threading: one context per host thread; calls on a context are serialized
 * on its HIP streams (a main stream, and a second one on which
 * edg_cluster_green runs each c chain beside its c+ chain).  Every call
 * returns an edg_status (0 = ok).
 */
#ifndef EDGPU_H
#define EDGPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    EDG_OK = 0,
    EDG_ERR_ADMISSION = 2, /* request exceeds the device-memory budget        */
    EDG_ERR_ARG = 3,       /* bad sector, size or argument                     */
    EDG_ERR_SOLVER = 4,    /* tridiagonal eigensolve failed (dstevx info /= 0) */
    EDG_ERR_STATE = 5,     /* no ground state held / wrong call order          */
    EDG_ERR_HIP = -1,      /* HIP runtime or launch error                      */
    EDG_ERR_NOMEM = -2     /* device allocation failed                         */
} edg_status;

typedef enum { EDG_FP64 = 0, EDG_FP32 = 1 } edg_precision;

/* Module globals of the reference made explicit (dmft_params.f90:78-188,
 * impurity_hamiltonian.f90:4-6, lanczos_solver.f90:4, ground_state.f90:7-8,
 * dmft_green.f90:5-6). */
typedef struct edg_params {
    int norb;          /* Norb                                                */
    int nbath;         /* Nbath  (nsite = norb + nbath <= 15)                 */
    int nspin;         /* Nspin (1 or 2)                                      */
    double U;          /* U                                                   */
    double J;          /* J (Jex)                                             */
    double Up;         /* Up; < 0 -> U - 2J  (dmft_params.f90:90-92)          */
    double Jp;         /* Jp; < 0 -> J       (dmft_params.f90:94-96)          */
    double mu;         /* Mu, on the impurity only                            */
    int maxnstep;      /* LanczosSteps                                        */
    int maxnstep_green;/* steps of each Green's-function chain; 0 = maxnstep
                          (the reference uses LanczosSteps for both)         */
    int diag_solver;   /* DiagSolver: 1 = Lanczos; 2 (ARPACK) -> EDG_ERR_ARG  */
    int precision;     /* edg_precision of the stored Lanczos vectors         */
    int both_spins;    /* 1: also run spin-down chains when spin-symmetric    */
    unsigned long long seed; /* start-vector seed (counter-based U[0,1))      */
} edg_params;

typedef struct edg_ctx edg_ctx;

/* Fill p with the reference's defaults (dmft_params.f90:78-188). */
void edg_default_params(edg_params *p);

/* Create a solver context on HIP device `device`. */
int edg_create(const edg_params *p, int device, edg_ctx **out);
int edg_destroy(edg_ctx *ctx);
/* Human-readable message of the last failing call on ctx (never NULL);
 * with ctx = NULL, that of the calling thread's last failing
 * context-free call (edg_hx_noalloc). */
const char *edg_last_error(const edg_ctx *ctx);
/* Device-memory budget for admission (bytes); 0 = no limit. */
int edg_set_memory_budget(edg_ctx *ctx, unsigned long long bytes);
/* Number of HIP devices visible to the process (no compute). */
int edg_device_count(int *n);
/* Kernel schedule of the H.v kernel: sector columns per workgroup (1, 2, 4
 * or 8, one workgroup per column group) or 16, 32, 64 for persistent
 * workgroups (one per compute unit, single-orbital hop table resident in
 * LDS): 16 = 1 column and 512 threads, 32 = 1 column and 1024 threads,
 * 64 = 2 columns and 1024 threads; and FP64 accumulation of the row sums (1) or accumulation in
 * the storage precision (0).  The two-orbital kernel takes one column for 1
 * and two for 2, 4 or 8 (when their LDS fits); 16, 32 and 64 send two-orbital
 * models to the general kernel.  Defaults: 2 columns for one orbital and for
 * two orbitals with FP32 vectors, else 1; FP64 row sums for FP32 vectors
 * with three or more orbitals.  No
 * counterpart in the reference (its hx is one scalar loop); set by the
 * autotuner or a caller that has measured. */
int edg_set_tuning(edg_ctx *ctx, int cols_per_block, int fp64_accumulate);
/* Form of the Lanczos step (no counterpart in the reference, whose steps
 * are the three-term recurrence of lanczos_solver.f90:147-170 in separate
 * vector passes).  on = 0 (default): an H.v pass plus an axpy / norm pass
 * per step.  on = 1: for one orbital with a one-workgroup-per-column-group
 * schedule, a single H.v pass per step that gathers from the previous
 * step's shifted residual (H - alpha_{j-2}) q_{j-1} - beta_{j-1} q_{j-2},
 * forms the new Krylov vector and the next residual in the same pass, and
 * takes alpha and beta from three FP64 sums reduced by a one-workgroup
 * finisher; it holds two more vectors per chain.  Both run the same
 * recurrence and differ in rounding only.  Measured on MI355X at Ns=14
 * (FP32), the fused form saves 14% per chain step alone but loses on the
 * whole solve (62 against 55 ms), whose two concurrent chains overlap the
 * two-pass form's H.v and axpy passes. */
int edg_set_fused_steps(edg_ctx *ctx, int on);
/* FP64 reorthogonalization of the Green's-function chains (the north star's
 * option; the reference runs the plain three-term recurrence,
 * lanczos_solver.f90:147-170, called at dmft_green.f90:218,248).  mode 0
 * (default): none.  mode 1: partial reorthogonalization, Simon's omega
 * recurrence (Math. Comp. 42 (1984) 115) estimating |<q_{j+1},q_k>| on the
 * device each step; when an estimate exceeds sqrt(eps) of the vector
 * storage, that step and the next are orthogonalized against every stored
 * chain vector by classical Gram-Schmidt with FP64 dot products and
 * updates.  mode 2: every step.  Each chain then keeps its whole Krylov
 * basis on the device (maxnstep_green + 1 vectors) instead of three.  The
 * ground-state Lanczos run is not affected. */
int edg_set_reorthogonalization(edg_ctx *ctx, int mode);
/* Steps of the last edg_cluster_green that were reorthogonalized, and steps
 * that ran the test (mode 1) or reorthogonalized (mode 2). */
int edg_reorth_stats(edg_ctx *ctx, long long *events, long long *steps);

/* Per-spin basis and hop tables of one (spin, N) sector half, built on the
 * device: ed_basis.f90:42-114 (generate_basis), the hybridization part of
 * impurity_hamiltonian.f90:146-174 with signs of :251-260, and the one-spin
 * diagonal of :119-141,232-239.  Out arrays (n = C(nsite,ne) entries,
 * kmax = norb*nbath hop slots, slot-major [k*n + r]):
 *   pattern[n]        bit pattern of rank r (0-based, increasing order)
 *   hop_col[kmax*n]   0-based target rank of hop k of rank r (-1 if unused)
 *   hop_val[kmax*n]   matrix element <r|H|target> (sign * V)
 *   hop_cnt[n]        number of used hop slots
 *   energy[n]         one-spin diagonal energy
 * Pass NULL for arrays not wanted; *n and *kmax are always filled. */
int edg_spin_table(edg_ctx *ctx, const double *em, const double *ek, const double *vmk,
                   int ispin, int ne, int *n, int *kmax, unsigned *pattern, int *hop_col,
                   double *hop_val, int *hop_cnt, double *energy);

/* hx (impurity_hamiltonian.f90:77-247): y = H x (py = 0) or y += H x
 * (py = 1) on sector (ne_up, ne_dn).  Host FP64 vectors; computed on the
 * device in the context's precision. */
int edg_hx(edg_ctx *ctx, const double *em, const double *ek, const double *vmk, int ne_up,
           int ne_dn, const double *x, double *y, int py);

/* edg_hx on device memory that allocates nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Needs no context: p gives the model (precision, norb, nbath, U, J, Up,
 * Jp, mu as for edg_create); em, ek, vmk are DEVICE arrays in the reference
 * layout (FP64, column-major); x and y are DEVICE sector vectors of
 * C(nsite,ne_up)*C(nsite,ne_dn) elements in p->precision (float or
 * double), not overlapping; y = H x (py = 0) or y += H x (py = 1).  The
 * per-spin tables go to a static device array of 16 MiB per GPU, held until
 * the process ends (every sector with norb <= 2 and nsite <= 15 fits; larger
 * tables are refused with EDG_ERR_ARG).  The kernels are queued on `stream`
 * (a hipStream_t; NULL = the null stream) on the current device and the call
 * returns without waiting; it calls no hipMalloc, hipFree, memset or copy,
 * so it may be captured in a HIP graph.  Calls that overlap on one GPU share
 * its array and race: serialize them.  Kernels and results are those of
 * edg_hx with a new context's default schedule.  Messages of failing calls:
 * edg_last_error(NULL), per host thread. */
int edg_hx_noalloc(const edg_params *p, const double *em, const double *ek, const double *vmk, int ne_up,
                   int ne_dn, const void *x, void *y, int py, void *stream);

/* lanczos_iteration (lanczos_solver.f90:92-173): plain 3-term Lanczos from
 * an unnormalized v_init for maxnstep steps.  a, b have maxnstep entries,
 * zero beyond *nstep; b[0] = 0.  Stops early when b(k) < 1e-16. */
int edg_lanczos_iteration(edg_ctx *ctx, const double *em, const double *ek, const double *vmk,
                          int ne_up, int ne_dn, const double *v_init, int *nstep, double *a,
                          double *b);

/* ground_state (ground_state.f90:1-81) with diag_lanczos
 * (lanczos_solver.f90:20-90): sweep `nsector` sectors given as
 * sectors[2*i] = ne_up, sectors[2*i+1] = ne_dn, keep the strict minimum.
 * The ground-state vector (gs%vec) stays on the device in ctx.
 * Outputs: *isector (1-based), *e0.  Sector i starts from the seeded
 * vector of seed + i - 1. */
int edg_ground_state(edg_ctx *ctx, const double *em, const double *ek, const double *vmk,
                     int nsector, const int *sectors, int *isector, double *e0);

/* Details of the held ground state: sector, dimension, Ritz value of the
 * tridiagonal (the reference's E0), Rayleigh quotient <gs|H|gs> accumulated
 * in FP64 (reported E0), residual |H gs - E0 gs|^2 (lanczos_solver.f90:293-297)
 * and Lanczos step count. */
int edg_gs_info(edg_ctx *ctx, int *ne_up, int *ne_dn, long long *dim, double *e0_ritz,
                double *e0, double *residual, int *nstep);
/* Copy gs%vec to the host (FP64, dim entries). */
int edg_get_gs_vector(edg_ctx *ctx, double *vec, long long dim);
/* Install a ground state (sector, E0, vector) given by the caller, e.g. one
 * computed elsewhere, so that cluster_green / observables can run on it. */
int edg_set_gs(edg_ctx *ctx, int ne_up, int ne_dn, double e0, const double *vec, long long dim);

/* cluster_green_ftn (dmft_green.f90:47-126) with green_diag (:177-260):
 * G_cl(nw,norb,nspin) complex (interleaved), omega[nw] Matsubara
 * frequencies.  coeffs (optional, may be NULL): ng x 4 x 2 x norb with
 * ng = maxnstep_green (or maxnstep)
 * FP64, column-major (i, {ap,bp,am,bm}, ispin, iorb) as written to
 * g_coeffs.dat (:89-109). */
int edg_cluster_green(edg_ctx *ctx, const double *em, const double *ek, const double *vmk,
                      int nw, const double *omega, double *G_cl, double *coeffs);

/* Occupancy n(iorb,ispin) (norb x 2, column-major) and double occupancy
 * docc(norb) of the held ground state (post_process.f90:63-112). */
int edg_observables(edg_ctx *ctx, double *occ, double *docc);

/* Self-energy and quasiparticle weight of post_process.f90 (dump_sigma
 * :121-185, Z :40-58), computed on the device from G_cl(nw,norb,nspin)
 * (interleaved complex, as returned by edg_cluster_green) and the bath:
 *   Sigma(iw,iorb,ispin) = i w + mu - em - D_cl - 1/G_cl,
 *   D_cl = sum_l vmk^2/(i w - ek) (dmft_green.f90:24-45),
 *   Z(iorb,ispin) = 1/(1 - Im Sigma(1)/omega(1)) (norb x nspin; may be NULL).
 * Every orbital's Sigma is kept: the reference's dump_sigma clears the whole
 * array inside its orbital loop (:161), so its Z is 1 for all but the last
 * orbital when norb > 1. */
int edg_self_energy(edg_ctx *ctx, const double *em, const double *ek, const double *vmk, int nw,
                    const double *omega, const double *G_cl, double *Sigma, double *Z);

/* apply_c (ed_operator.f90:13-76): vec_out = c+_{iorb,ispin} vec (pm = 1)
 * or c_{iorb,ispin} vec (pm = 2), from sector (ne_up, ne_dn) to the sector
 * with one more / one fewer electron of spin ispin. */
int edg_apply_c(edg_ctx *ctx, int ne_up, int ne_dn, const double *vec, int pm, int iorb,
                int ispin, double *vec_out);

/* Timing of the H.v kernel on device-resident vectors of sector
 * (ne_up, ne_dn) in the context's precision: `reps` launches bracketed by
 * HIP events on the context's stream after `warmup` launches.  Returns the
 * average milliseconds per launch and the algorithmic bytes of one launch
 * (x read once + y written once = 2 * s * dim). */
int edg_time_hv(edg_ctx *ctx, const double *em, const double *ek, const double *vmk, int ne_up,
                int ne_dn, int warmup, int reps, double *ms_per_launch, double *bytes_per_launch);

/* Streaming reference (no counterpart in the reference): average
 * milliseconds of a copy y = x (4n bytes read, 4n written) and of an add
 * y = x + z (8n read, 4n written) over n floats on the context's stream,
 * hand-written 16-byte-access kernels timed with HIP events over `reps`
 * launches after 5 warm-up ones: the plain-streaming rate the H.v's
 * fabric traffic is measured against. */
int edg_time_stream(edg_ctx *ctx, long long n, int reps, double *copy_ms, double *add_ms);

/* Full solve timing helper: ground state in one sector + all Green's
 * function chains, returning wall milliseconds of each phase measured with
 * HIP events on the context stream (phase_ms[0] ground state,
 * phase_ms[1] chains) and the number of H.v applications performed. */
int edg_last_solve_stats(edg_ctx *ctx, double *phase_ms, long long *n_hv, long long *n_steps);

/* ---------------------------------------------------------------------
 * Batched solver and request loop.  No counterpart in the reference (it
 * solves one impurity problem per call); for callers with many independent
 * one- or two-orbital problems: the candidate sectors and parameter points
 * of a DMFT step or scan, the c+/c chains of every orbital and spin, or a
 * stream of requests.  Every problem runs the same
 * algorithm as edg_ground_state + edg_cluster_green (stored-basis Lanczos
 * ground state, FP64 Rayleigh-quotient E0, c+/c chains, continued
 * fraction); the Lanczos steps of all problems in flight share batched
 * launches (one H.v, one axpy and one step-advance kernel per step), and
 * problems are admitted by device bytes as earlier ones finish.  A
 * problem's result does not depend on what else is in the batch.
 * --------------------------------------------------------------------- */
typedef struct edg_batch edg_batch;

/* One impurity problem: norb = 1 (or 0) a single orbital with em(1,2) in
 * em[]; norb = 2 the two-orbital Kanamori impurity with em(2,2) in em2[]
 * and the couplings J, Up, Jp (Up < 0: U - 2J, Jp < 0: J, as
 * dmft_params.f90:90-96).  Bath of nbath sites with ek(nbath,2) and
 * vmk(norb,nbath,2) column-major as for edg_ground_state, U and mu, in the
 * sector (nup, ndn). */
typedef struct {
    int nbath;
    double U;
    double mu;
    double em[2];
    const double *ek;
    const double *vmk;
    int nup;
    int ndn;
    int maxnstep;        /* ground-state Lanczos steps */
    int maxnstep_green;  /* steps of each c+/c chain (0: maxnstep) */
    int both_spins;      /* also run the spin-down chains when spin symmetry would skip them */
    unsigned long long seed;
    int gs_only;         /* 1: ground state only (E0, no chains, G = 0), e.g. for a sector sweep */
    int norb;            /* impurity orbitals: 0 or 1, or 2 */
    double J, Up, Jp;    /* norb = 2: Hund coupling, U' and pair hopping */
    double em2[4];       /* norb = 2: em(2,2) column-major */
} edg_problem;

/* precision: EDG_FP64 / EDG_FP32 vector storage; nspin 1 or 2 (G_cl spins);
 * omega: the nw Matsubara frequencies; budget_bytes: device bytes the
 * problems in flight may hold (0 = 85% of the free device memory, counting
 * as free the idle chunks of earlier queues); a problem is also held back
 * while the device cannot allocate it.  Problem memory is sub-allocated
 * from large device chunks that are not unmapped while the process runs,
 * and the pinned host areas of the problems from 16 MiB slabs;
 * edg_batch_destroy leaves a queue's chunks and slabs (up to 1 GiB of
 * them) to later queues, and edg_batch_release_memory returns the idle
 * ones. */
int edg_batch_create(int device, int precision, int nspin, int nw, const double *omega,
                     unsigned long long budget_bytes, edg_batch **out);
int edg_batch_destroy(edg_batch *b);
/* Return to the device the idle memory chunks that destroyed queues left
 * on it, and free their idle pinned host slabs (no counterpart in the
 * reference). */
int edg_batch_release_memory(int device);
const char *edg_batch_last_error(const edg_batch *b);
/* Replay the batched steps as captured HIP graphs (1, default) or launch
 * their kernels one by one (0).  With graphs, a chunk of steps is captured
 * once the set of chains in flight has stayed unchanged for two chunks and
 * replayed until it changes; the chunks in between launch directly. */
int edg_batch_set_graphs(edg_batch *b, int use_graphs);
/* H.v tile schedule by size class (set while the engine is idle): entry i
 * applies to every chain of problems with nsite <= nsite_max[i] (increasing);
 * wide[i] = 1 gives one sector column per 512-thread tile, wide[i] = 0 gives
 * clamp(tile_elems[i] / nup, 1, 64) columns of nup rows per 256-thread tile.
 * The default is one class: tiles of about 2048 elements, one column per
 * 512-thread tile from nup >= 1024.  Set by the autotuner (edgpu.autotune). */
int edg_batch_set_schedule(edg_batch *b, int n, const int *nsite_max, const int *tile_elems, const int *wide);
/* Device bytes one problem holds while in flight (its admission size). */
int edg_batch_problem_bytes(const edg_batch *b, const edg_problem *p, unsigned long long *bytes);
/* Queue a problem; *id receives its handle.  EDG_ERR_ADMISSION when it can
 * never fit the budget, EDG_ERR_ARG for an invalid problem. */
int edg_batch_submit(edg_batch *b, const edg_problem *p, long long *id);
/* Admit queued problems while the budget allows, run the per-problem stages
 * that are ready and up to max_steps batched Lanczos steps.  *pending
 * receives the number of submitted problems not yet finished.  Admission
 * follows submission order by device bytes, except that a problem whose
 * state vector takes 32 MiB or more waits while two other such problems
 * run, and smaller problems behind it are admitted meanwhile. */
int edg_batch_step(edg_batch *b, int max_steps, int *pending);
/* Run until every submitted problem is finished. */
int edg_batch_run(edg_batch *b);
/* Pop one finished problem: returns 1 and fills id, e0 (FP64 Rayleigh
 * quotient), G (G_cl(nw,norb,nspin) as interleaved re/im), nstep
 * (ground-state Lanczos steps) and coeffs (ng x 4 x 2 x norb, ng =
 * maxnstep_green or maxnstep: (ap, bp, am, bm) per spin and orbital as
 * written to g_coeffs.dat, dmft_green.f90:89-109; the layout of
 * edg_cluster_green), or returns 0 when none is ready.  Pointers may be
 * NULL. */
int edg_batch_poll(edg_batch *b, long long *id, double *e0, double *G, int *nstep, double *coeffs);
/* stats[0..11] (the caller passes 12 doubles): [0] batched steps,
 * [1] chain-steps, [2] problems finished, [3] device bytes in use,
 * [4] budget, [5] problems in flight, [6] graph launches, [7] descriptor
 * rebuilds; host seconds inside edg_batch_step spent on [8] admission
 * (tables, start vectors), [9] stage transitions (read-backs, tridiagonal
 * eigenvectors, combine, apply_c, continued fractions), [10] queueing the
 * steps, [11] blocked on device events. */
int edg_batch_stats(const edg_batch *b, double *stats);
/* Graph lifecycle counters (no counterpart in the reference): graph execs
 * instantiated and destroyed so far, and execs alive now (current plus
 * retired ones whose last launch the stream has not yet passed).  A chunk
 * is captured only after the chain set has stayed unchanged for two
 * chunks, so instantiations stay below the descriptor rebuilds. */
int edg_batch_graph_stats(const edg_batch *b, long long *instantiated, long long *destroyed, int *live);

#ifdef __cplusplus
}
#endif
#endif /* EDGPU_H */
