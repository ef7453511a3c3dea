/*
 * mpm_hip.h -- C ABI of libmpm_hip.so, the MI355X (gfx950) implementation of
 * MaterialPointSolver.jl's explicit single-phase MUSL timestep.
 *
 * The library replaces the reference's device plugin for this path
 * (ext/AMDExt/devicehelpfunc_amd.jl plus the KernelAbstractions kernels that
 * procedure!(..., Val(:OS), Val(:MUSL), Val(:a|:v)) launches).  Each entry
 * point below names the reference interface it stands in for.  A Julia
 * package extension binds these with ccall (a sketch ships in the package's
 * julia/ folder); Python binds them with ctypes.
 *
 * Conventions
 *  - Every call returns MPM_OK (0) or a negative MPM_E_* code; the message of
 *    the last failure is kept per context (mpm_last_error).
 *  - Host arrays are the reference's Julia layouts: column-major (rows, k)
 *    matrices, i.e. k contiguous streams of `rows` values, in FP64 / Int64.
 *    Indices (nid, BC flags) keep the reference's 1-based meaning.
 *  - Grid-node arrays (BC) use the reference's node order: 2D rows from the
 *    top (y descending, src/types/grid.jl:92-100), 3D y fastest then x then z
 *    (src/types/grid.jl:236-249).
 *  - The library never keeps a host pointer after a call returns.
 *  - A context is driven from one thread; contexts are independent.
 *  - Step calls are asynchronous on the context's own HIP stream; download,
 *    mpm_sync and mpm_destroy synchronise.
 */
#ifndef MPM_HIP_H
#define MPM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPM_ABI_VERSION 2

enum {
    MPM_OK = 0,
    MPM_E_ARG = -1,         /* invalid argument or configuration */
    MPM_E_UNSUPPORTED = -2, /* a switch value this library does not implement */
    MPM_E_HIP = -3,         /* HIP runtime error (message carries hipGetErrorString) */
    MPM_E_NOMEM = -4,       /* device allocation failed */
    MPM_E_STATE = -5,       /* call out of order (e.g. step before upload) */
    MPM_E_RANGE = -6        /* a particle left the grid or became non-finite */
};

enum { MPM_FP64 = 0, MPM_FP32 = 1 };                 /* precision */
enum { MPM_UGIMP = 0, MPM_LINEAR = 1 };               /* basis */
enum { MPM_MUSL = 0, MPM_USL = 1, MPM_USF = 2 };     /* scheme */
enum { MPM_VA_A = 0, MPM_VA_V = 1 };                  /* va */
enum { MPM_LINEARELASTIC = 0, MPM_DRUCKERPRAGER = 1, MPM_MOHRCOULOMB = 2 }; /* constitutive */

/* Device-side model description, fixed for the life of a context: the
 * scalars of Grid2D/Grid3D and Particle2D/Particle3D the step reads
 * (src/types/grid.jl:33-75,170-217; src/types/particle.jl:33-81,160-210) plus
 * the precision chosen by the structs' T2 (the reference's ϵ keyword). */
typedef struct mpm_config {
    int32_t dim;           /* 2 | 3 */
    int32_t precision;     /* MPM_FP64: every field FP64 (parity mode).
                              MPM_FP32: FP32 bulk arithmetic and grid atomics,
                              FP64 particle positions and det F. */
    int32_t basis;         /* MPM_UGIMP (grid.NIC 9 | 27) or MPM_LINEAR (NIC 4 | 8) */
    int32_t sort_interval; /* steps between on-device cell sorts; 0 = never */
    int64_t np;            /* number of particles (mp.np) */
    int64_t nmat;          /* number of materials (length of the property vectors) */
    int64_t nnx, nny, nnz; /* nodes per axis (grid.nnx, nny, nnz; nnz = 1 in 2D) */
    double x1, y1, z1;     /* grid origin (grid.x1, y1, z1) */
    double dx, dy, dz;     /* grid spacing (grid.dx, dy, dz) */
    double lpx, lpy, lpz;  /* particle spacing (mp.dx, dy, dz), uGIMP domain size */
} mpm_config;

/* Per-call switches: the fields of Args2D/Args3D that procedure! reads
 * (src/types/modelargs.jl:31-58,122-149; src/solvers/OS_MUSL.jl:12-102,
 * OS_USL.jl:102-144,325-366, OS_USF.jl:11-97). */
typedef struct mpm_args {
    int32_t scheme;        /* args.scheme: MPM_MUSL, MPM_USL or MPM_USF.  A USF step updates
                              the stress first, with the previous step's nodal displacement;
                              its position update is committed by the next step's stress
                              update (downloads add it), so the state after n USF steps is
                              the reference's.  The first USF step of a context (or after
                              steps of another scheme) applies the nodal displacement held
                              on the grid: zero on a fresh context, as in the reference. */
    int32_t va;            /* MPM_VA_A | MPM_VA_V (args.va) */
    int32_t constitutive;  /* MPM_LINEARELASTIC | MPM_DRUCKERPRAGER | MPM_MOHRCOULOMB
                              (3D Mohr-Coulomb is defined by this library: a principal-space
                              return with dilation psi; the reference's 3D mcP! is a no-op,
                              src/materials/mohrcoulomb.jl:152-455) */
    int32_t mvl;           /* args.MVL: cell-averaged mean stress */
    double flip, pic;      /* args.FLIP, args.PIC */
    double zeta_s;         /* args.ζs, solid damping */
    double gravity;        /* args.gravity, on the last axis */
    double Te;             /* args.Te: gravity ramp end and plasticity gate */
    double gravity_x;      /* body force on the first axis (no reference counterpart: the
                              reference's gravity acts on the last axis only; 0 reproduces it).
                              For slope-aligned frames; ramped like gravity. */
    double gravity_y;      /* 3D: body force on the second axis (ignored in 2D); 0 by default */
} mpm_args;

typedef struct mpm_ctx mpm_ctx;

/* ABI version of the loaded library (MPM_ABI_VERSION). */
int mpm_abi_version(void);

/* Create a context on the current HIP device: validates cfg, allocates every
 * device buffer the step needs and zeroes the grid.
 * Replaces the allocation half of host2device(grid, mp, attr, bc, Val(:ROCm))
 * (ext/AMDExt/devicehelpfunc_amd.jl:17-37). */
int mpm_create(const mpm_config *cfg, mpm_ctx **out);

/* Upload particle state (np rows each).  xi, vs: np x dim; sigma, eps_s:
 * np x (4|6) in the reference's Voigt order ([xx,yy,zz,xy] 2D,
 * [xx,yy,zz,xy,yz,xz] 3D); F: np x (4|9) row-major tensor per particle;
 * omega, omega0, rho, rho0, eps_q, eps_k: np; nid: np, 1-based material ids.
 * Replaces the particle half of host2device (devicehelpfunc_amd.jl:26). */
int mpm_upload_particles(mpm_ctx *ctx, const double *xi, const double *vs,
                         const double *sigma, const double *eps_s, const double *F,
                         const double *omega, const double *omega0, const double *rho,
                         const double *rho0, const double *eps_q, const double *eps_k,
                         const int64_t *nid);

/* Upload the material table (nmat entries each; attr.Ks, Gs, ϕ, ψ, c, σt, Hp, cr,
 * ϕr; src/types/property.jl:31-52).  Replaces host2device for attr
 * (devicehelpfunc_amd.jl:27).  ϕr (may be NULL = all zero) is read by the 3D
 * Mohr-Coulomb return only: with Hp < 0, c > cr and ϕr > 0, friction softens
 * linearly from ϕ to ϕr and dilation from ψ to 0 over the plastic strain in
 * which cohesion softens from c to cr (mohr_coulomb.tcc:8-25); the reference's
 * 2D mcP! softens cohesion only (mohrcoulomb.jl:36), and so does this library
 * in 2D. */
int mpm_upload_materials(mpm_ctx *ctx, const double *Ks, const double *Gs,
                         const double *phi, const double *psi, const double *c,
                         const double *sigma_t, const double *Hp, const double *cr,
                         const double *phir);

/* Upload the solid velocity boundary (bc.v{x,y,z}_s_idx / _val, ni entries each in
 * the reference's node order; vz_* ignored in 2D; src/types/boundary.jl:115-131,169-190).
 * Replaces host2device for bc (devicehelpfunc_amd.jl:28). */
int mpm_upload_bc(mpm_ctx *ctx, const int64_t *vx_idx, const double *vx_val,
                  const int64_t *vy_idx, const double *vy_val,
                  const int64_t *vz_idx, const double *vz_val);

/* One timestep at time Ti: the whole of
 * procedure!(args, grid, mp, attr, bc, ΔT, Ti, Val(:OS), Val(:MUSL), Val(args.va))
 * (src/solvers/OS_MUSL.jl:12-102): gravity ramp Ti < Te, plasticity when
 * Ti >= Te.  Asynchronous. */
int mpm_step(mpm_ctx *ctx, const mpm_args *args, double dt, double Ti);

/* nsteps timesteps with the submit_work! clock: Ti += dt after each step,
 * accumulated in FP64 (src/solver.jl:118-124).  *Ti_out receives the final
 * time (may be NULL).  Asynchronous. */
int mpm_run(mpm_ctx *ctx, const mpm_args *args, double dt, double Ti, int64_t nsteps,
            double *Ti_out);

/* mpm_run with the steps between two sorts captured into HIP graphs and
 * replayed: runs of two or more steps whose kernel arguments do not change
 * (Ti >= Te) become one graph launch; sorts and gravity-ramp steps stay eager.
 * Graphs are cached per context (run length, arguments, dt, buffer parity).
 * Same results as mpm_run up to the order of the atomic sums. */
int mpm_run_graph(mpm_ctx *ctx, const mpm_args *args, double dt, double Ti, int64_t nsteps,
                  double *Ti_out);

/* Runtime options of a context (no reference counterpart; for the autotuner
 * and the benchmark):
 *  MPM_OPT_SORT_INTERVAL: steps between sorts, >= 1; only for contexts created
 *                         with sort_interval > 0 (the sort buffers exist).
 *  MPM_OPT_TILES:         1 = LDS tile kernels after a sort (default), 0 = the
 *                         plain global-atomic kernels.
 *  MPM_OPT_PROFILE_MASK:  bit k set = mpm_profile_enable times kernel k of
 *                         mpm_profile_kernel_name's order; 0 = every kernel
 *                         (default).  Fewer event pairs per step keep the
 *                         profiled steps closer to unprofiled ones.
 *  MPM_OPT_SCATTER:       form of the tile scatter of P2G #1 (and of the MUSL
 *                         re-map's, see MPM_OPT_REMAP_SCATTER); -1 (default):
 *                         MPM_SCATTER_PAIRS in FP32 3D uGIMP, elsewhere
 *                         MPM_SCATTER_ROUNDS:
 *                         MPM_SCATTER_ROUNDS: rank rounds of distinct bases
 *                         adding into an LDS window;
 *                         MPM_SCATTER_OWNED: each lane owns one stencil base
 *                         of a tile and sums its particles in registers (3D
 *                         P2G: mass and momentum, then force);
 *                         MPM_SCATTER_OWNED_LAYERS: the same, one z layer of
 *                         the stencil per pass (fewer registers, more passes);
 *                         MPM_SCATTER_ROUNDS_LAYERED: rank rounds of one
 *                         group adding into three LDS windows, one per
 *                         stencil z offset, so each round's LDS updates form
 *                         three independent chains (FP32 3D uGIMP; elsewhere
 *                         it runs as MPM_SCATTER_ROUNDS);
 *                         MPM_SCATTER_PAIRS: pair rounds: a lane takes two
 *                         particles of one stencil base and does one LDS
 *                         read-add-write per slot for both (FP32 3D uGIMP;
 *                         elsewhere it runs as MPM_SCATTER_OWNED).
 *  MPM_OPT_FUSED_REMAP:   0 (default): with the tile kernels, the MUSL re-map
 *                         is a gather kernel, then the momentum scatter;
 *                         1: it gathers, moves the particles and scatters the
 *                         new momentum in one kernel (rank rounds: every
 *                         precision; the base-owned forms: not in FP64 3D
 *                         uGIMP, whose sums exceed the register file).
 *  MPM_OPT_TILE_PARTIALS: 1 (default): the tile scatters store each tile's
 *                         window sums and the grid solves add the windows that
 *                         hold a node, in a fixed order (no flush atomics);
 *                         0: the windows are flushed with global atomics.
 *  MPM_OPT_REMAP_SCATTER: form of the MUSL re-map's momentum scatter (an
 *                         MPM_SCATTER_* value); -1 (default): base-owned
 *                         (MPM_SCATTER_OWNED) in FP32 3D uGIMP, elsewhere the
 *                         form P2G #1 runs in. */
enum { MPM_OPT_SORT_INTERVAL = 0, MPM_OPT_TILES = 1, MPM_OPT_PROFILE_MASK = 2, MPM_OPT_SCATTER = 3,
       MPM_OPT_FUSED_REMAP = 4, MPM_OPT_TILE_PARTIALS = 5, MPM_OPT_REMAP_SCATTER = 6 };
enum { MPM_SCATTER_ROUNDS = 0, MPM_SCATTER_OWNED = 1, MPM_SCATTER_OWNED_LAYERS = 2,
       MPM_SCATTER_ROUNDS_LAYERED = 3, MPM_SCATTER_PAIRS = 4 };
int mpm_set_option(mpm_ctx *ctx, int option, int64_t value);

/* Wait for all work on the context's stream; reports range/finite flags the
 * kernels raised.  Replaces KAsync(getBackend(Val(:ROCm))) (src/solver.jl:128). */
int mpm_sync(mpm_ctx *ctx);

/* Download the device2host! field set, in the caller's original particle order
 * (ext/AMDExt/devicehelpfunc_amd.jl:39-69): xi, vs (np x dim), sigma, eps_s
 * (np x (4|6)), sigma_m, omega, eps_q, eps_k, eps_v, ms (np).  Any pointer may
 * be NULL to skip that field.  Synchronises. */
int mpm_download_particles(mpm_ctx *ctx, double *xi, double *vs, double *sigma,
                           double *sigma_m, double *omega, double *eps_q, double *eps_k,
                           double *eps_v, double *eps_s, double *ms);

/* Free every device buffer of the context and the context itself.
 * Replaces clean_device!(..., Val(:ROCm)) (devicehelpfunc_amd.jl:71-102). */
void mpm_destroy(mpm_ctx *ctx);

/* Message of the last failed call on ctx (or of the last failed mpm_create
 * when ctx is NULL).  Never NULL. */
const char *mpm_last_error(const mpm_ctx *ctx);

/* Per-kernel timing with HIP events on the context's stream.  Enable before
 * the steps to measure; mpm_profile_read fills up to n entries of the kernel
 * names' order given by mpm_profile_kernel_name: total milliseconds and
 * launch counts since the last enable.  Synchronises.
 * No reference counterpart (the reference's perf() reports only it/s,
 * src/toolkits/terminaltxt.jl:67-85). */
int mpm_profile_enable(mpm_ctx *ctx, int on);
int mpm_profile_read(mpm_ctx *ctx, double *total_ms, int64_t *launches, int n);
const char *mpm_profile_kernel_name(int k);
int mpm_profile_kernel_count(void);

/* Number of grid nodes that receive a non-zero basis weight from at least one
 * particle at the current positions (the "active" nodes whose traffic the
 * step's algorithmic byte count includes).  Diagnostic; synchronises.
 * No reference counterpart. */
int mpm_count_active_nodes(mpm_ctx *ctx, int64_t *count);

/* mpm_count_active_nodes without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  No reference counterpart.  The node
 * bitmap (one bit per node) and the counter live in one file-scope device
 * array of 4 MiB + 8 B that each GPU keeps until the process ends; grids of
 * more than 33,554,432 nodes are rejected with MPM_E_RANGE.  Every call on a
 * GPU shares that array (it is cleared at the start of each call): calls that
 * overlap on one GPU, from any context or thread, race. */
int mpm_count_active_nodes_noalloc(mpm_ctx *ctx, int64_t *count);

/* Triad bandwidth probe A = B + C on nbytes-per-array FP32 buffers, in GB/s
 * (10^9 B/s), best of `reps`.  Replaces Tpeak(Val(:ROCm))
 * (ext/AMDExt/devicehelpfunc_amd.jl:104-142), which reports GiB/s of the same
 * triad. */
int mpm_tpeak(int64_t nbytes, int reps, double *gbps);

/* ---------------------------------------------------------------------------
 * Learned return map (no reference counterpart: the reference has no
 * surrogate).  A 16-128-128-128-7 SiLU MLP run on the matrix cores in BF16
 * with FP32 accumulation.  Inputs per sample (16): the trial stress as its
 * principal values p1 >= p2 >= p3 over the scale S = max|p| + c cos(phi) (3)
 * and its principal frame as a unit quaternion (x, y, z; w >= 0) (3), the
 * strain increment [xx,yy,zz,xy,yz,xz] (engineering shears) times 2 G / S (6),
 * the accumulated plastic strain eps_q, the friction angle (rad), c / S and the
 * dilation angle (rad), with c, phi, psi the strength in effect.  Outputs (7):
 * scores of the four branches of the principal-space Mohr-Coulomb return
 * (main plane, edge s1 = s2, edge s2 = s3, apex) and three unused outputs
 * (trained to zero).  In a step (mpm_set_surrogate) a yielding particle takes
 * the highest-scoring branch's closed form when that branch's conditions and
 * the yield residual check hold, else the exact return; the return-mapping
 * log (mpm_return_log) records these inputs with the exact branch one-hot.
 * Parameters: one FP32 vector of mpm_mlp_param_count() values, W1[128][16] b1
 * W2[128][128] b2 W3[128][128] b3 W4[7][128] b4 followed by x_mean[16]
 * x_istd[16] y_scale[7] (inputs are standardised, outputs are scaled).
 * Samples are row-major (n x 16 inputs, n x 7 outputs), FP32.
 * ------------------------------------------------------------------------- */
typedef struct mpm_mlp mpm_mlp;

int mpm_mlp_param_count(void);
/* batch_capacity: largest minibatch / gradient batch (32 .. 4194304). */
int mpm_mlp_create(int64_t batch_capacity, mpm_mlp **out);
void mpm_mlp_destroy(mpm_mlp *m);
const char *mpm_mlp_last_error(const mpm_mlp *m);
int mpm_mlp_set_params(mpm_mlp *m, const float *params, int64_t n);
int mpm_mlp_get_params(mpm_mlp *m, float *params, int64_t n);
/* y = network(x) in output units, n rows (the fused matrix-core kernel). */
int mpm_mlp_forward(mpm_mlp *m, const float *x, int64_t n, float *y);
/* Mean squared error over the normalised outputs of n <= capacity samples and
 * its gradient (first mpm_mlp_param_count() - 39 entries: the trainable ones). */
int mpm_mlp_grad(mpm_mlp *m, const float *x, const float *y, int64_t n, float *grad,
                 float *loss);
/* `steps` Adam steps (lr, beta 0.9 / 0.999) on minibatches of `batch` samples
 * drawn with replacement (seeded); loss_history (may be NULL) gets each step's
 * minibatch loss. */
int mpm_mlp_train(mpm_mlp *m, const float *x, const float *y, int64_t n, int64_t steps,
                  int64_t batch, double lr, uint64_t seed, float *loss_history);
/* Mean duration of the forward kernel on n synthetic samples (HIP events). */
int mpm_mlp_bench(mpm_mlp *m, int64_t n, int reps, double *ms_per_launch);

/* Attach a trained network as the 3D Mohr-Coulomb return of this context (FP32
 * mode, 3D, no MVL; NULL detaches).  Each plastic step then stores the trial
 * stress with its principal values and frame (G2P #2), evaluates the network
 * for every particle, and returns a yielding particle on the branch the
 * network scores highest when that branch's conditions hold and the returned
 * stress satisfies the yield function to |f| <= tol (2 c cos(phi) +
 * |s1 + s3| sin(phi)); every other yielding particle takes the exact return.
 * The network must outlive the attachment.  Resets the counters below. */
int mpm_set_surrogate(mpm_ctx *ctx, mpm_mlp *m, double tol);
/* Yielding particle-steps and fallbacks to the exact return since attaching. */
int mpm_surrogate_stats(mpm_ctx *ctx, int64_t *plastic, int64_t *fallbacks);
/* Log return-mapping samples of the exact 3D Mohr-Coulomb return: every
 * stride-th particle (in the current particle order) that yields, up to
 * `capacity` samples (0 disables and frees); a sample is the network's 16
 * inputs and the exact return's branch as a one-hot over the first four
 * outputs (the last three zero).  Resets the log. */
int mpm_return_log(mpm_ctx *ctx, int64_t capacity, int64_t stride);
/* Copy up to max_rows logged samples (x: rows x 16, y: rows x 7, row-major). */
int mpm_return_log_read(mpm_ctx *ctx, float *x, float *y, int64_t max_rows, int64_t *rows);

#ifdef __cplusplus
}
#endif

#endif /* MPM_HIP_H */
