/*
 * mpm_oracle.c -- FP64 CPU restatement of MaterialPointSolver.jl's single-phase
 * explicit MPM step (OS coupling; MUSL, USL and USF schemes).  TEST
 * INFRASTRUCTURE ONLY.
 *
 * This file is the parity checker for the HIP library.  Only tests, the smoke
 * check and the CPU-baseline leg of bench.py may load it; the product path
 * never calls it.
 *
 * Data layout follows the reference exactly: every per-particle array is the
 * Julia column-major (np, k) matrix, i.e. k contiguous streams of np values,
 * and every index (p2c, p2n, nid, BC flags) is 1-based.  The basis values are
 * stored per particle in NIC slots exactly like the reference's Nij / dN
 * arrays, filled in candidate order from the reference's p2nD tables.
 *
 * Each function names the reference kernel it restates (paths relative to the
 * reference repository root).
 *
 * Threading: nthreads == 1 runs every loop serially in index order (fully
 * deterministic).  nthreads > 1 runs the particle and node loops with OpenMP
 * and turns the scatters into atomic adds, which is what the reference's
 * KernelAbstractions CPU backend does (Atomix atomics); summation order is then
 * free, as in the reference.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct oracle_model {
    /* ---- switches (src/types/modelargs.jl) ---- */
    int32_t dim;          /* 2 or 3 */
    int32_t nic;          /* 9 / 4 (2D), 27 / 8 (3D) */
    int32_t basis;        /* 0 = uGIMP, 1 = linear */
    int32_t va;           /* 0 = :a, 1 = :v */
    int32_t constitutive; /* 0 = linearelastic, 1 = druckerprager, 2 = mohrcoulomb */
    int32_t mvl;          /* 0 / 1 */
    int32_t nthreads;     /* 1 = serial deterministic */
    int32_t pad0;
    double flip, pic, zeta_s, gravity, Te;
    /* ---- grid (src/types/grid.jl) ---- */
    int64_t nnx, nny, nnz, ncx, ncy, ncz, ni, nc;
    double x1, y1, z1, dx, dy, dz;
    const double *gxi;     /* ni x dim node coordinates, reference node order */
    const int64_t *p2nD;   /* candidate table, rows x (2|4) column-major */
    int64_t p2nD_rows;
    double *g_ms, *g_ps, *g_fs, *g_vs, *g_vsT, *g_as, *g_dus; /* ni, ni x dim */
    double *g_sigm, *g_omega;                                  /* nc */
    /* ---- particles (src/types/particle.jl) ---- */
    int64_t np;
    double lpx, lpy, lpz;  /* mp.dx, mp.dy, mp.dz */
    double *xi, *vs, *ps;  /* np x dim */
    double *sig, *sij, *eps_s, *deps_s; /* np x (4|6) */
    double *dF, *F;        /* np x (4|9) */
    double *sigm, *eps_q, *eps_k, *eps_v, *omega, *omega0, *ms, *rho, *rho0; /* np */
    int64_t *p2c;          /* np */
    int64_t *p2n;          /* np x nic */
    double *Nij, *dNx, *dNy, *dNz; /* np x nic */
    /* ---- property (src/types/property.jl) ---- */
    const int64_t *nid;    /* np, 1-based */
    const double *Ks, *Gs, *c, *phi, *psi, *sigt, *Hp, *cr; /* nmat */
    const double *phir;    /* nmat: residual friction (3D Mohr-Coulomb softening) */
    /* ---- velocity boundary (src/types/boundary.jl) ---- */
    const int64_t *vx_idx, *vy_idx, *vz_idx; /* ni, != 0 -> fixed */
    const double *vx_val, *vy_val, *vz_val;
    /* ---- extension: body force on the other axes (0 = the reference) ---- */
    double gravity_x, gravity_y;
    double Ti_now;         /* time of the current step (set by oracle_procedure) */
    int32_t scheme;        /* 0 = :MUSL, 1 = :USL, 2 = :USF */
    int32_t pad1;
} oracle_model;

#define COL(a, n, i, k) ((a)[(int64_t)(k) * (n) + (i)])   /* (i,k) 0-based of (n, .) */

static inline int par(const oracle_model *m) { return m->nthreads > 1; }

/* ---------------------------------------------------------------------------
 * Julia scalar semantics used by the reference.
 * ------------------------------------------------------------------------- */

/* sign(x): -1, 0 or 1 with the sign of zero kept (Base.sign). */
static inline double jl_sign(double x) {
    if (x > 0.0) return 1.0;
    if (x < 0.0) return -1.0;
    return x; /* +-0.0, NaN */
}

/* Base.min / Base.max for IEEE floats: NaN propagates, -0.0 < +0.0. */
static inline double jl_min(double x, double y) {
    if (isnan(x)) return x;
    if (isnan(y)) return y;
    if (y < x || (signbit(y) && !signbit(x))) return y;
    return x;
}
static inline double jl_max(double x, double y) {
    if (isnan(x)) return x;
    if (isnan(y)) return y;
    if (y > x || (!signbit(y) && signbit(x))) return y;
    return x;
}

/* fld(x, y) / cld(x, y) for floats: floor / ceiling of the EXACT quotient
 * (the Julia manual's example: fld(6.0, 0.1) == 59.0 although 6.0/0.1 == 60.0).
 * The candidate quotient is corrected with the exact remainder x - q*y, which
 * fma evaluates with a single rounding and which is exact here because
 * |x - q*y| < 2|y| and both terms share an exponent range. */
double oracle_fld(double x, double y) {
    double q = floor(x / y);
    double r = fma(-q, y, x);
    if (y > 0.0) {
        if (r < 0.0) q -= 1.0;
        else if (r >= y) q += 1.0;
    } else {
        if (r > 0.0) q -= 1.0;
        else if (r <= y) q += 1.0;
    }
    return q;
}
double oracle_cld(double x, double y) {
    double q = ceil(x / y);
    double r = fma(-q, y, x);
    if (y > 0.0) {
        if (r > 0.0) q += 1.0;
        else if (r <= -y) q -= 1.0;
    } else {
        if (r < 0.0) q += 1.0;
        else if (r >= -y) q -= 1.0;
    }
    return q;
}

/* ---------------------------------------------------------------------------
 * Basis functions: src/toolkits/mpbasisfunc.jl
 * ------------------------------------------------------------------------- */

/* linearBasis (mpbasisfunc.jl:24-29) */
void oracle_linear_basis(double d, double h, double *N, double *dN) {
    double hd = 1.0 / h;
    *N = 1.0 - fabs(d) * hd;
    *dN = -jl_sign(d) * hd;
}

/* uGIMPbasis (mpbasisfunc.jl:40-53) -- the CPU path's direct form. */
void oracle_ugimp_basis(double d, double h, double lp, double *N, double *dN) {
    double Ni = 0.0, dNi = 0.0, a = fabs(d);
    if (a < 0.5 * lp) {
        Ni = 1.0 - ((4.0 * d * d + lp * lp) / (4.0 * h * lp));
        dNi = -((8.0 * d) / (4.0 * h * lp));
    } else if ((0.5 * lp) <= a && a < (h - 0.5 * lp)) {
        Ni = 1.0 - (a / h);
        dNi = jl_sign(d) * (-1.0 / h);
    } else if ((h - 0.5 * lp) <= a && a < (h + 0.5 * lp)) {
        Ni = ((h + 0.5 * lp - a) * (h + 0.5 * lp - a)) / (2.0 * h * lp);
        dNi = -jl_sign(d) * ((h + 0.5 * lp - a) / (h * lp));
    }
    *N = Ni;
    *dN = dNi;
}

/* ---------------------------------------------------------------------------
 * Cell and candidate-node indices: utils_OS.jl:203-205, 325-328 and
 * src/toolkits/p2nindex.jl.
 * ------------------------------------------------------------------------- */

/* p2c (1-based) as in resetmpstatus_OS_CPU! (utils_OS.jl:256-258, 384-387).
 * The reference sums the float fld/cld results and truncates. */
int64_t oracle_p2c(const oracle_model *m, double x, double y, double z) {
    double s;
    if (m->dim == 2) {
        s = oracle_cld(y - m->y1, m->dy) + oracle_fld(x - m->x1, m->dx) * (double)m->ncy;
    } else {
        s = oracle_cld(y - m->y1, m->dy) +
            oracle_fld(z - m->z1, m->dz) * (double)m->ncy * (double)m->ncx +
            oracle_fld(x - m->x1, m->dx) * (double)m->ncy;
    }
    return (int64_t)s;
}

/* Julia integer cld for positive divisor. */
static inline int64_t icld(int64_t a, int64_t b) {
    int64_t q = a / b;
    if ((a % b != 0) && ((a > 0) == (b > 0))) q += 1;
    return q;
}

/* getP2N_uGIMP / getP2N_linear (p2nindex.jl:35-45, 76-88, 130-140, 243-255).
 * iy is the 1-based candidate row; the result is a 1-based node id. */
int64_t oracle_p2n(const oracle_model *m, int64_t p2c, int64_t iy) {
    const int64_t R = m->p2nD_rows;
    const int64_t *t = m->p2nD;
    int64_t r = iy - 1;
    if (m->dim == 2) {
        int64_t col_id = icld(p2c, m->ncy);
        int64_t row_id = p2c - (col_id - 1) * m->ncy;
        int64_t ltn_id = col_id * m->nny - row_id;
        return ltn_id + t[r] + t[R + r];
    } else {
        int64_t layer = icld(p2c, m->ncx * m->ncy);
        int64_t l_idx = icld(p2c - (layer - 1) * m->ncx * m->ncy, m->ncy);
        int64_t l_idy = (p2c - (layer - 1) * m->ncx * m->ncy) - (l_idx - 1) * m->ncy;
        if (m->basis == 0) {
            return (layer + t[r]) * m->nnx * m->nny + (l_idx + t[R + r]) * m->nny +
                   (l_idy + t[2 * R + r]) + t[3 * R + r];
        }
        return (layer + t[r]) * m->nnx * m->nny + (l_idx + t[R + r]) * m->nny +
               (l_idy + t[2 * R + r]) + 1;
    }
}

/* ---------------------------------------------------------------------------
 * resetgridstatus_OS! (utils_OS.jl:58-99)
 * ------------------------------------------------------------------------- */
void oracle_resetgrid(oracle_model *m) {
    const int64_t ni = m->ni, D = m->dim;
    #pragma omp parallel for if (par(m)) schedule(static)
    for (int64_t i = 0; i < ni; ++i) {
        if (i < m->nc) {
            m->g_sigm[i] = 0.0;
            m->g_omega[i] = 0.0;
        }
        m->g_ms[i] = 0.0;
        for (int64_t d = 0; d < D; ++d) {
            COL(m->g_ps, ni, i, d) = 0.0;
            COL(m->g_fs, ni, i, d) = 0.0;
        }
    }
}

/* ---------------------------------------------------------------------------
 * resetmpstatus_OS_CPU! (uGIMP; utils_OS.jl:235-277 2D, 363-415 3D) and
 * resetmpstatus_OS! (linear; utils_OS.jl:101-166).  Also the body of
 * initmpstatus! (src/solvers/extras.jl:91-179) when update_mass == 0.
 * ------------------------------------------------------------------------- */
static void reset_particle(oracle_model *m, int64_t ix, int update_mass) {
    const int64_t np = m->np, D = m->dim, NIC = m->nic, ni = m->ni;
    double px = COL(m->xi, np, ix, 0), py = COL(m->xi, np, ix, 1);
    double pz = D == 3 ? COL(m->xi, np, ix, 2) : 0.0;
    if (update_mass) {
        double mpms = m->omega[ix] * m->rho[ix];
        m->ms[ix] = mpms;
        for (int64_t d = 0; d < D; ++d) COL(m->ps, np, ix, d) = mpms * COL(m->vs, np, ix, d);
    }
    int64_t p2c = oracle_p2c(m, px, py, pz);
    m->p2c[ix] = p2c;
    if (m->basis == 1) {
        /* linear: every one of the NIC cell corners, no support test */
        for (int64_t iy = 1; iy <= NIC; ++iy) {
            int64_t p2n = oracle_p2n(m, p2c, iy);
            COL(m->p2n, np, ix, iy - 1) = p2n;
            double Nx, dNx, Ny, dNy, Nz = 1.0, dNz = 0.0;
            oracle_linear_basis(px - COL(m->gxi, ni, p2n - 1, 0), m->dx, &Nx, &dNx);
            oracle_linear_basis(py - COL(m->gxi, ni, p2n - 1, 1), m->dy, &Ny, &dNy);
            if (D == 3) {
                oracle_linear_basis(pz - COL(m->gxi, ni, p2n - 1, 2), m->dz, &Nz, &dNz);
                COL(m->Nij, np, ix, iy - 1) = Nx * Ny * Nz;
                COL(m->dNx, np, ix, iy - 1) = dNx * Ny * Nz;
                COL(m->dNy, np, ix, iy - 1) = dNy * Nx * Nz;
                COL(m->dNz, np, ix, iy - 1) = dNz * Nx * Ny;
            } else {
                COL(m->Nij, np, ix, iy - 1) = Nx * Ny;
                COL(m->dNx, np, ix, iy - 1) = dNx * Ny;
                COL(m->dNy, np, ix, iy - 1) = dNy * Nx;
            }
        }
        return;
    }
    for (int64_t iy = 0; iy < NIC; ++iy) {
        COL(m->Nij, np, ix, iy) = 0.0;
        COL(m->dNx, np, ix, iy) = 0.0;
        COL(m->dNy, np, ix, iy) = 0.0;
        if (D == 3) COL(m->dNz, np, ix, iy) = 0.0;
    }
    int64_t viy = 1;
    const int64_t ncand = D == 2 ? 16 : 64;
    for (int64_t iy = 1; iy <= ncand; ++iy) {
        int64_t p2n = oracle_p2n(m, p2c, iy);
        double ddx = px - COL(m->gxi, ni, p2n - 1, 0);
        double ddy = py - COL(m->gxi, ni, p2n - 1, 1);
        double ddz = D == 3 ? pz - COL(m->gxi, ni, p2n - 1, 2) : 0.0;
        int inside = fabs(ddx) < (m->dx + 0.5 * m->lpx) && fabs(ddy) < (m->dy + 0.5 * m->lpy);
        if (D == 3) inside = inside && fabs(ddz) < (m->dz + 0.5 * m->lpz);
        if (inside) {
            double Nx, dNx, Ny, dNy;
            oracle_ugimp_basis(ddx, m->dx, m->lpx, &Nx, &dNx);
            oracle_ugimp_basis(ddy, m->dy, m->lpy, &Ny, &dNy);
            if (D == 3) {
                double Nz, dNz;
                oracle_ugimp_basis(ddz, m->dz, m->lpz, &Nz, &dNz);
                COL(m->Nij, np, ix, viy - 1) = Nx * Ny * Nz;
                COL(m->dNx, np, ix, viy - 1) = dNx * Ny * Nz;
                COL(m->dNy, np, ix, viy - 1) = dNy * Nx * Nz;
                COL(m->dNz, np, ix, viy - 1) = dNz * Nx * Ny;
            } else {
                COL(m->Nij, np, ix, viy - 1) = Nx * Ny;
                COL(m->dNx, np, ix, viy - 1) = dNx * Ny;
                COL(m->dNy, np, ix, viy - 1) = dNy * Nx;
            }
            COL(m->p2n, np, ix, viy - 1) = p2n;
            viy += 1;
        }
        if (viy > NIC) break;
    }
}

void oracle_resetmp(oracle_model *m) {
    #pragma omp parallel for if (par(m)) schedule(static)
    for (int64_t ix = 0; ix < m->np; ++ix) reset_particle(m, ix, 1);
}

/* initmpstatus! (extras.jl:91-179): topology and basis only, run once. */
void oracle_initmp(oracle_model *m) {
    #pragma omp parallel for if (par(m)) schedule(static)
    for (int64_t ix = 0; ix < m->np; ++ix) reset_particle(m, ix, 0);
}

static inline void scatter_add(const oracle_model *m, double *a, double v) {
    if (par(m)) {
        #pragma omp atomic
        *a += v;
    } else {
        *a += v;
    }
}

/* ---------------------------------------------------------------------------
 * P2G_OS! (utils_OS.jl:424-452 2D, 461-496 3D)
 * ------------------------------------------------------------------------- */
void oracle_p2g(oracle_model *m, double gravity) {
    /* body force on the other axes (an extension of the reference, zero there) */
    const double Te = m->Te, Ti = m->Ti_now;
    const double gx = Ti < Te ? m->gravity_x / Te * Ti : m->gravity_x;
    const double gy = Ti < Te ? m->gravity_y / Te * Ti : m->gravity_y;
    const int64_t np = m->np, ni = m->ni, NIC = m->nic, D = m->dim;
    #pragma omp parallel for if (par(m)) schedule(static)
    for (int64_t ix = 0; ix < np; ++ix) {
        for (int64_t iy = 0; iy < NIC; ++iy) {
            double Ni = COL(m->Nij, np, ix, iy);
            if (Ni != 0.0) {
                double dNx = COL(m->dNx, np, ix, iy), dNy = COL(m->dNy, np, ix, iy);
                int64_t n = COL(m->p2n, np, ix, iy) - 1;
                double vol = m->omega[ix];
                double NiM = m->ms[ix] * Ni;
                scatter_add(m, &m->g_ms[n], NiM);
                scatter_add(m, &COL(m->g_ps, ni, n, 0), Ni * COL(m->ps, np, ix, 0));
                scatter_add(m, &COL(m->g_ps, ni, n, 1), Ni * COL(m->ps, np, ix, 1));
                if (D == 2) {
                    double fx = -vol * (dNx * COL(m->sig, np, ix, 0) + dNy * COL(m->sig, np, ix, 3));
                    if (gx != 0.0) fx += NiM * gx;
                    scatter_add(m, &COL(m->g_fs, ni, n, 0), fx);
                    scatter_add(m, &COL(m->g_fs, ni, n, 1),
                                -vol * (dNy * COL(m->sig, np, ix, 1) + dNx * COL(m->sig, np, ix, 3)) +
                                    NiM * gravity);
                } else {
                    double dNz = COL(m->dNz, np, ix, iy);
                    scatter_add(m, &COL(m->g_ps, ni, n, 2), Ni * COL(m->ps, np, ix, 2));
                    double fx = -vol * (dNx * COL(m->sig, np, ix, 0) + dNy * COL(m->sig, np, ix, 3) +
                                        dNz * COL(m->sig, np, ix, 5));
                    double fy = -vol * (dNy * COL(m->sig, np, ix, 1) + dNx * COL(m->sig, np, ix, 3) +
                                        dNz * COL(m->sig, np, ix, 4));
                    if (gx != 0.0) fx += NiM * gx;
                    if (gy != 0.0) fy += NiM * gy;
                    scatter_add(m, &COL(m->g_fs, ni, n, 0), fx);
                    scatter_add(m, &COL(m->g_fs, ni, n, 1), fy);
                    scatter_add(m, &COL(m->g_fs, ni, n, 2),
                                -vol * (dNz * COL(m->sig, np, ix, 2) + dNx * COL(m->sig, np, ix, 5) +
                                        dNy * COL(m->sig, np, ix, 4)) +
                                    NiM * gravity);
                }
            }
        }
    }
}

static inline int64_t bc_idx(const int64_t *a, int64_t i) { return a[i]; }

/* ---------------------------------------------------------------------------
 * solvegrid_a_OS! (utils_OS.jl:547-583 2D, 595-640 3D)
 * solvegrid_OS!   (utils_OS.jl:507-535 2D, 650-687 3D)
 * USL / USF: solvegrid_a_USL_OS! (OS_USL.jl:147-187 2D, 225-269 3D) and
 * solvegrid_USL_OS! (OS_USL.jl:23-56 2D, 63-100 3D) are the same arithmetic
 * plus the nodal displacement dus = v_new * dt, and leave ps alone.  In the :a
 * form the reference stores the new velocity in grid.vs; it is kept in g_vsT
 * here (the pre-update velocity is not read again in those schemes).
 * ------------------------------------------------------------------------- */
void oracle_solvegrid(oracle_model *m, double dt) {
    const int64_t ni = m->ni, D = m->dim;
    const double eps = DBL_EPSILON;
    const int64_t *idx[3] = {m->vx_idx, m->vy_idx, m->vz_idx};
    const double *val[3] = {m->vx_val, m->vy_val, m->vz_val};
    #pragma omp parallel for if (par(m)) schedule(static)
    for (int64_t i = 0; i < ni; ++i) {
        double msd = m->g_ms[i] < eps ? 0.0 : 1.0 / m->g_ms[i];
        double Fs[3];
        if (m->va == 0) {
            for (int64_t d = 0; d < D; ++d)
                if (bc_idx(idx[d], i) != 0) COL(m->g_ps, ni, i, d) = val[d][i];
            for (int64_t d = 0; d < D; ++d) COL(m->g_vs, ni, i, d) = COL(m->g_ps, ni, i, d) * msd;
            double s = 0.0;
            for (int64_t d = 0; d < D; ++d) s += COL(m->g_fs, ni, i, d) * COL(m->g_fs, ni, i, d);
            double dampvs = -m->zeta_s * sqrt(s);
            for (int64_t d = 0; d < D; ++d)
                Fs[d] = COL(m->g_fs, ni, i, d) + dampvs * jl_sign(COL(m->g_vs, ni, i, d));
            for (int64_t d = 0; d < D; ++d)
                COL(m->g_vsT, ni, i, d) = COL(m->g_vs, ni, i, d) + Fs[d] * dt * msd;
            for (int64_t d = 0; d < D; ++d) COL(m->g_as, ni, i, d) = Fs[d] * msd;
            for (int64_t d = 0; d < D; ++d)
                if (bc_idx(idx[d], i) != 0) COL(m->g_vsT, ni, i, d) = val[d][i];
            for (int64_t d = 0; d < D; ++d)
                if (bc_idx(idx[d], i) != 0) COL(m->g_as, ni, i, d) = val[d][i];
        } else {
            double P[3];
            for (int64_t d = 0; d < D; ++d) P[d] = COL(m->g_ps, ni, i, d);
            for (int64_t d = 0; d < D; ++d) COL(m->g_vs, ni, i, d) = P[d] * msd;
            double s = 0.0;
            for (int64_t d = 0; d < D; ++d) s += COL(m->g_fs, ni, i, d) * COL(m->g_fs, ni, i, d);
            double dampvs = -m->zeta_s * sqrt(s);
            for (int64_t d = 0; d < D; ++d)
                Fs[d] = COL(m->g_fs, ni, i, d) + dampvs * jl_sign(COL(m->g_vs, ni, i, d));
            for (int64_t d = 0; d < D; ++d) COL(m->g_vsT, ni, i, d) = (P[d] + Fs[d] * dt) * msd;
            for (int64_t d = 0; d < D; ++d)
                if (bc_idx(idx[d], i) != 0) COL(m->g_vsT, ni, i, d) = val[d][i];
        }
        if (m->scheme == 0) {
            for (int64_t d = 0; d < D; ++d) COL(m->g_ps, ni, i, d) = 0.0;
        } else {
            for (int64_t d = 0; d < D; ++d) COL(m->g_dus, ni, i, d) = COL(m->g_vsT, ni, i, d) * dt;
        }
    }
}

/* ---------------------------------------------------------------------------
 * doublemapping1_a_OS! (utils_OS.jl:745-788 2D, 856-908 3D)
 * doublemapping1_OS!   (utils_OS.jl:698-743 2D, 799-854 3D)
 * doublemapping1_a_USL_OS! (OS_USL.jl:189-223 2D, 271-313 3D): as the :a form
 * but the PIC term gathers the NEW nodal velocity (the USL grid solve
 * overwrote grid.vs with it).  USL/USF :v use doublemapping1_OS! unchanged.
 * The reference loads Ks/Gs there without using them; nothing to restate.
 * ------------------------------------------------------------------------- */
void oracle_dm1(oracle_model *m, double dt) {
    const int64_t np = m->np, ni = m->ni, NIC = m->nic, D = m->dim;
    #pragma omp parallel for if (par(m)) schedule(static)
    for (int64_t ix = 0; ix < np; ++ix) {
        double a[3] = {0, 0, 0}, b[3] = {0, 0, 0}, c[3] = {0, 0, 0};
        for (int64_t iy = 0; iy < NIC; ++iy) {
            double Ni = COL(m->Nij, np, ix, iy);
            if (Ni != 0.0) {
                int64_t n = COL(m->p2n, np, ix, iy) - 1;
                for (int64_t d = 0; d < D; ++d) {
                    double vT = COL(m->g_vsT, ni, n, d);
                    if (m->va == 0) {
                        a[d] += Ni * vT;                          /* tmp_p */
                        b[d] += Ni * COL(m->g_as, ni, n, d);      /* tmp_v */
                        c[d] += Ni * (m->scheme == 0 ? COL(m->g_vs, ni, n, d) : vT);
                    } else {
                        a[d] += Ni * vT;                          /* tmp_pos */
                        b[d] += Ni * (vT - COL(m->g_vs, ni, n, d)); /* tmp_s1 */
                        c[d] += Ni * vT;                          /* tmp_s2 */
                    }
                }
            }
        }
        for (int64_t d = 0; d < D; ++d) {
            if (m->va == 0) {
                COL(m->xi, np, ix, d) += a[d] * dt;
                COL(m->vs, np, ix, d) = m->flip * COL(m->vs, np, ix, d) + m->pic * c[d] + b[d] * dt;
            } else {
                COL(m->xi, np, ix, d) += dt * a[d];
                COL(m->vs, np, ix, d) = m->flip * (COL(m->vs, np, ix, d) + b[d]) + m->pic * c[d];
            }
            COL(m->ps, np, ix, d) = m->ms[ix] * COL(m->vs, np, ix, d);
        }
    }
}

/* doublemapping2_OS! (utils_OS.jl:917-959) */
void oracle_dm2(oracle_model *m) {
    const int64_t np = m->np, ni = m->ni, NIC = m->nic, D = m->dim;
    #pragma omp parallel for if (par(m)) schedule(static)
    for (int64_t ix = 0; ix < np; ++ix) {
        for (int64_t iy = 0; iy < NIC; ++iy) {
            double Ni = COL(m->Nij, np, ix, iy);
            if (Ni != 0.0) {
                int64_t n = COL(m->p2n, np, ix, iy) - 1;
                for (int64_t d = 0; d < D; ++d)
                    scatter_add(m, &COL(m->g_ps, ni, n, d), COL(m->ps, np, ix, d) * Ni);
            }
        }
    }
}

/* doublemapping3_OS! (utils_OS.jl:968-1016) */
void oracle_dm3(oracle_model *m, double dt) {
    const int64_t ni = m->ni, D = m->dim;
    const double eps = DBL_EPSILON;
    const int64_t *idx[3] = {m->vx_idx, m->vy_idx, m->vz_idx};
    const double *val[3] = {m->vx_val, m->vy_val, m->vz_val};
    #pragma omp parallel for if (par(m)) schedule(static)
    for (int64_t i = 0; i < ni; ++i) {
        double msd = m->g_ms[i] < eps ? 0.0 : 1.0 / m->g_ms[i];
        for (int64_t d = 0; d < D; ++d) COL(m->g_vs, ni, i, d) = COL(m->g_ps, ni, i, d) * msd;
        for (int64_t d = 0; d < D; ++d)
            if (bc_idx(idx[d], i) != 0) COL(m->g_vs, ni, i, d) = val[d][i];
        for (int64_t d = 0; d < D; ++d) COL(m->g_dus, ni, i, d) = COL(m->g_vs, ni, i, d) * dt;
    }
}

/* ---------------------------------------------------------------------------
 * G2P_OS! (utils_OS.jl:1025-1074 2D, 1083-1153 3D)
 * ------------------------------------------------------------------------- */
void oracle_g2p(oracle_model *m, double dt) {
    const int64_t np = m->np, ni = m->ni, NIC = m->nic, D = m->dim;
    #pragma omp parallel for if (par(m)) schedule(static)
    for (int64_t ix = 0; ix < np; ++ix) {
        if (D == 2) {
            double r = 1.0 / dt; /* inv(ΔT) */
            double dF1 = 0, dF2 = 0, dF3 = 0, dF4 = 0;
            for (int64_t iy = 0; iy < NIC; ++iy) {
                if (COL(m->Nij, np, ix, iy) != 0.0) {
                    int64_t n = COL(m->p2n, np, ix, iy) - 1;
                    double dNx = COL(m->dNx, np, ix, iy), dNy = COL(m->dNy, np, ix, iy);
                    dF1 += COL(m->g_dus, ni, n, 0) * dNx;
                    dF2 += COL(m->g_dus, ni, n, 0) * dNy;
                    dF3 += COL(m->g_dus, ni, n, 1) * dNx;
                    dF4 += COL(m->g_dus, ni, n, 1) * dNy;
                }
            }
            COL(m->dF, np, ix, 0) = dF1; COL(m->dF, np, ix, 1) = dF2;
            COL(m->dF, np, ix, 2) = dF3; COL(m->dF, np, ix, 3) = dF4;
            double exx = dF1 * r, eyy = dF4 * r, exy = 0.5 * (dF2 + dF3) * r;
            m->eps_v[ix] = sqrt(exx * exx + eyy * eyy + 2.0 * exy * exy);
            COL(m->deps_s, np, ix, 0) = dF1;
            COL(m->deps_s, np, ix, 1) = dF4;
            COL(m->deps_s, np, ix, 3) = dF2 + dF3;
            COL(m->eps_s, np, ix, 0) += dF1;
            COL(m->eps_s, np, ix, 1) += dF4;
            COL(m->eps_s, np, ix, 3) += dF2 + dF3;
            double F1 = COL(m->F, np, ix, 0), F2 = COL(m->F, np, ix, 1);
            double F3 = COL(m->F, np, ix, 2), F4 = COL(m->F, np, ix, 3);
            COL(m->F, np, ix, 0) = (dF1 + 1.0) * F1 + dF2 * F3;
            COL(m->F, np, ix, 1) = (dF1 + 1.0) * F2 + dF2 * F4;
            COL(m->F, np, ix, 2) = (dF4 + 1.0) * F3 + dF3 * F1;
            COL(m->F, np, ix, 3) = (dF4 + 1.0) * F4 + dF3 * F2;
            double J = COL(m->F, np, ix, 0) * COL(m->F, np, ix, 3) -
                       COL(m->F, np, ix, 1) * COL(m->F, np, ix, 2);
            m->omega[ix] = J * m->omega0[ix];
            m->rho[ix] = m->rho0[ix] / J;
        } else {
            double r = 1.0 / dt;
            double f[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int64_t iy = 0; iy < NIC; ++iy) {
                if (COL(m->Nij, np, ix, iy) != 0.0) {
                    int64_t n = COL(m->p2n, np, ix, iy) - 1;
                    double dNx = COL(m->dNx, np, ix, iy), ds1 = COL(m->g_dus, ni, n, 0);
                    double dNy = COL(m->dNy, np, ix, iy), ds2 = COL(m->g_dus, ni, n, 1);
                    double dNz = COL(m->dNz, np, ix, iy), ds3 = COL(m->g_dus, ni, n, 2);
                    f[0] += ds1 * dNx; f[1] += ds1 * dNy; f[2] += ds1 * dNz;
                    f[3] += ds2 * dNx; f[4] += ds2 * dNy; f[5] += ds2 * dNz;
                    f[6] += ds3 * dNx; f[7] += ds3 * dNy; f[8] += ds3 * dNz;
                }
            }
            for (int k = 0; k < 9; ++k) COL(m->dF, np, ix, k) = f[k];
            double exx = f[0] * r, eyy = f[4] * r, ezz = f[8] * r;
            double exy = 0.5 * (f[1] + f[3]) * r;
            double eyz = 0.5 * (f[5] + f[7]) * r;
            double exz = 0.5 * (f[2] + f[6]) * r;
            m->eps_v[ix] = sqrt(exx * exx + eyy * eyy + ezz * ezz +
                                2.0 * (exy * exy + eyz * eyz + exz * exz));
            COL(m->deps_s, np, ix, 0) = f[0];
            COL(m->deps_s, np, ix, 1) = f[4];
            COL(m->deps_s, np, ix, 2) = f[8];
            COL(m->deps_s, np, ix, 3) = f[1] + f[3];
            COL(m->deps_s, np, ix, 4) = f[5] + f[7];
            COL(m->deps_s, np, ix, 5) = f[2] + f[6];
            COL(m->eps_s, np, ix, 0) += f[0];
            COL(m->eps_s, np, ix, 1) += f[4];
            COL(m->eps_s, np, ix, 2) += f[8];
            COL(m->eps_s, np, ix, 3) += f[1] + f[3];
            COL(m->eps_s, np, ix, 4) += f[5] + f[7];
            COL(m->eps_s, np, ix, 5) += f[2] + f[6];
            double F[9];
            for (int k = 0; k < 9; ++k) F[k] = COL(m->F, np, ix, k);
            double G[9];
            G[0] = (f[0] + 1.0) * F[0] + f[1] * F[3] + f[2] * F[6];
            G[1] = (f[0] + 1.0) * F[1] + f[1] * F[4] + f[2] * F[7];
            G[2] = (f[0] + 1.0) * F[2] + f[1] * F[5] + f[2] * F[8];
            G[3] = (f[4] + 1.0) * F[3] + f[3] * F[0] + f[5] * F[6];
            G[4] = (f[4] + 1.0) * F[4] + f[3] * F[1] + f[5] * F[7];
            G[5] = (f[4] + 1.0) * F[5] + f[3] * F[2] + f[5] * F[8];
            G[6] = (f[8] + 1.0) * F[6] + f[7] * F[3] + f[6] * F[0];
            G[7] = (f[8] + 1.0) * F[7] + f[7] * F[4] + f[6] * F[1];
            G[8] = (f[8] + 1.0) * F[8] + f[7] * F[5] + f[6] * F[2];
            for (int k = 0; k < 9; ++k) COL(m->F, np, ix, k) = G[k];
            double J = G[0] * G[4] * G[8] + G[1] * G[5] * G[6] + G[2] * G[3] * G[7] -
                       G[6] * G[4] * G[2] - G[7] * G[5] * G[0] - G[8] * G[3] * G[1];
            m->omega[ix] = J * m->omega0[ix];
            m->rho[ix] = m->rho0[ix] / J;
        }
    }
}

static inline void deviatoric(oracle_model *m, int64_t ix) {
    const int64_t np = m->np;
    double sm = (COL(m->sig, np, ix, 0) + COL(m->sig, np, ix, 1) + COL(m->sig, np, ix, 2)) * 0.333333;
    m->sigm[ix] = sm;
    COL(m->sij, np, ix, 0) = COL(m->sig, np, ix, 0) - sm;
    COL(m->sij, np, ix, 1) = COL(m->sig, np, ix, 1) - sm;
    COL(m->sij, np, ix, 2) = COL(m->sig, np, ix, 2) - sm;
    COL(m->sij, np, ix, 3) = COL(m->sig, np, ix, 3);
    if (m->dim == 3) {
        COL(m->sij, np, ix, 4) = COL(m->sig, np, ix, 4);
        COL(m->sij, np, ix, 5) = COL(m->sig, np, ix, 5);
    }
}

/* ---------------------------------------------------------------------------
 * liE! (src/materials/linearelastic.jl:22-69 2D, 78-140 3D)
 * ------------------------------------------------------------------------- */
void oracle_liE(oracle_model *m) {
    const int64_t np = m->np;
    #pragma omp parallel for if (par(m)) schedule(static)
    for (int64_t ix = 0; ix < np; ++ix) {
        int64_t k = m->nid[ix] - 1;
        double Ks = m->Ks[k], Gt = m->Gs[k];
        double *s = m->sig;
        if (m->dim == 2) {
            double wxy = 0.5 * (COL(m->dF, np, ix, 1) - COL(m->dF, np, ix, 2));
            double s1 = COL(s, np, ix, 0), s2 = COL(s, np, ix, 1), s4 = COL(s, np, ix, 3);
            COL(s, np, ix, 0) += wxy * s4 * 2.0;
            COL(s, np, ix, 1) += -wxy * s4 * 2.0;
            COL(s, np, ix, 3) += wxy * (s2 - s1);
            double Dt = Ks + 1.333333 * Gt, Dd = Ks - 0.666667 * Gt;
            double e1 = COL(m->deps_s, np, ix, 0), e2 = COL(m->deps_s, np, ix, 1);
            double e3 = COL(m->deps_s, np, ix, 2), e4 = COL(m->deps_s, np, ix, 3);
            COL(s, np, ix, 0) += Dt * e1 + Dd * e2 + Dd * e3;
            COL(s, np, ix, 1) += Dd * e1 + Dt * e2 + Dd * e3;
            COL(s, np, ix, 2) += Dd * e1 + Dd * e2 + Dt * e3;
            COL(s, np, ix, 3) += Gt * e4;
        } else {
            double wxy = 0.5 * (COL(m->dF, np, ix, 3) - COL(m->dF, np, ix, 1));
            double wyz = 0.5 * (COL(m->dF, np, ix, 7) - COL(m->dF, np, ix, 5));
            double wxz = 0.5 * (COL(m->dF, np, ix, 6) - COL(m->dF, np, ix, 2));
            double s1 = COL(s, np, ix, 0), s2 = COL(s, np, ix, 1), s3 = COL(s, np, ix, 2);
            double s4 = COL(s, np, ix, 3), s5 = COL(s, np, ix, 4), s6 = COL(s, np, ix, 5);
            COL(s, np, ix, 0) += 2.0 * (s4 * wxy + s6 * wxz);
            COL(s, np, ix, 1) += -2.0 * (s4 * wxy - s5 * wyz);
            COL(s, np, ix, 2) += -2.0 * (s6 * wxz + s5 * wyz);
            COL(s, np, ix, 3) += wxy * (s2 - s1) + wxz * s5 + wyz * s6;
            COL(s, np, ix, 4) += wyz * (s3 - s2) - wxz * s4 - wxy * s6;
            COL(s, np, ix, 5) += wxz * (s3 - s1) + wxy * s5 - wyz * s4;
            double Dt = Ks + 1.333333 * Gt, Dd = Ks - 0.666667 * Gt;
            double e1 = COL(m->deps_s, np, ix, 0), e2 = COL(m->deps_s, np, ix, 1);
            double e3 = COL(m->deps_s, np, ix, 2);
            COL(s, np, ix, 0) += Dt * e1 + Dd * e2 + Dd * e3;
            COL(s, np, ix, 1) += Dd * e1 + Dt * e2 + Dd * e3;
            COL(s, np, ix, 2) += Dd * e1 + Dd * e2 + Dt * e3;
            COL(s, np, ix, 3) += Gt * COL(m->deps_s, np, ix, 3);
            COL(s, np, ix, 4) += Gt * COL(m->deps_s, np, ix, 4);
            COL(s, np, ix, 5) += Gt * COL(m->deps_s, np, ix, 5);
        }
        deviatoric(m, ix);
    }
}

/* ---------------------------------------------------------------------------
 * dpP! (src/materials/druckerprager.jl:22-80 2D, 89-155 3D)
 * ------------------------------------------------------------------------- */
void oracle_dpP(oracle_model *m) {
    const int64_t np = m->np;
    const int ns = m->dim == 2 ? 4 : 6;
    #pragma omp parallel for if (par(m)) schedule(static)
    for (int64_t ix = 0; ix < np; ++ix) {
        double sm = m->sigm[ix];
        int64_t k = m->nid[ix] - 1;
        double c = m->c[k], phi = m->phi[k], psi = m->psi[k], st = m->sigt[k];
        double Gs = m->Gs[k], Ks = m->Ks[k];
        double *S = m->sij;
        double q = COL(S, np, ix, 0) * COL(S, np, ix, 0) + COL(S, np, ix, 1) * COL(S, np, ix, 1) +
                   COL(S, np, ix, 2) * COL(S, np, ix, 2);
        double tau;
        if (m->dim == 2) {
            tau = sqrt(0.5 * q + COL(S, np, ix, 3) * COL(S, np, ix, 3));
        } else {
            tau = sqrt(0.5 * q + COL(S, np, ix, 3) * COL(S, np, ix, 3) +
                       COL(S, np, ix, 4) * COL(S, np, ix, 4) + COL(S, np, ix, 5) * COL(S, np, ix, 5));
        }
        double kphi = (6.0 * c * cos(phi)) / (1.732051 * (3.0 + sin(phi)));
        double qphi = (6.0 * sin(phi)) / (1.732051 * (3.0 + sin(phi)));
        double qpsi = (6.0 * sin(psi)) / (1.732051 * (3.0 + sin(psi)));
        st = jl_min(st, kphi / qphi);
        double ab = sqrt(1.0 + qphi * qphi) - qphi;
        double tb = kphi - qphi * st;
        double fs = tau + qphi * sm - kphi;
        double ft = sm - st;
        double BF = (tau - tb) - (ab * (sm - st));
        if (((sm < st) && (fs > 0.0)) || ((sm >= st) && (BF > 0.0))) {
            double dls = fs / (Gs + Ks * qphi * qpsi);
            double t1 = sm - Ks * qpsi * dls;
            double t2 = (kphi - qphi * t1) / tau;
            COL(m->sig, np, ix, 0) = COL(S, np, ix, 0) * t2 + t1;
            COL(m->sig, np, ix, 1) = COL(S, np, ix, 1) * t2 + t1;
            COL(m->sig, np, ix, 2) = COL(S, np, ix, 2) * t2 + t1;
            for (int j = 3; j < ns; ++j) COL(m->sig, np, ix, j) = COL(S, np, ix, j) * t2;
            m->eps_q[ix] += dls * sqrt(0.333333 + 0.222222 * qpsi * qpsi);
            m->eps_k[ix] += dls * qpsi;
        }
        if ((sm >= st) && (BF <= 0.0)) {
            double dlt = ft / Ks;
            COL(m->sig, np, ix, 0) = COL(S, np, ix, 0) + st;
            COL(m->sig, np, ix, 1) = COL(S, np, ix, 1) + st;
            COL(m->sig, np, ix, 2) = COL(S, np, ix, 2) + st;
            m->eps_q[ix] += dlt * 0.333333 * 1.414214;
            m->eps_k[ix] += dlt;
        }
        deviatoric(m, ix);
    }
}

/* ---------------------------------------------------------------------------
 * 3D Mohr-Coulomb (build-defined).  The reference's 3D mcP! body is commented
 * out (src/materials/mohrcoulomb.jl:152-455), so there is no reference
 * arithmetic to restate; parity with the reference is unpinned.  This is the
 * FP64 definition the HIP kernel is checked against, pinned in
 * tests/test_oracle_mc3d.py by closed forms and a constrained minimisation.
 *
 * Same conventions as the reference's 2D mcP! (mohrcoulomb.jl:22-86): tension
 * positive, cohesion softening c = max(c + Hp eps_q, cr), apex at c cot(phi),
 * plastic strain increment from the stress correction through the inverse
 * elastic matrix (Dt = K + 1.333333 G, Dd = K - 0.666667 G), eps_k assigned its
 * trace, eps_q += sqrt(0.666667 (e_n . e_n + 2 e_s . e_s)).
 *
 * Return mapping in principal stresses for perfect plasticity with dilation
 * psi (de Souza Neto, Peric & Owen, Computational Methods for Plasticity,
 * 2008, sec. 8.2: one-vector return to the main plane, two-vector return to
 * an edge, return to the apex), with the planes written as
 *   f = (s1 - s3) + (s1 + s3) sin(phi) - 2 c cos(phi),  s1 >= s2 >= s3,
 *   flow direction N = (1 + sin(psi), 0, -(1 - sin(psi))).
 * ------------------------------------------------------------------------- */

/* Eigen-decomposition of a symmetric 3x3 tensor [xx,yy,zz,xy,yz,xz] by cyclic
 * Jacobi rotations (Golub & Van Loan, Matrix Computations, sec. 8.5).  lam:
 * eigenvalues in descending order; V: column k = eigenvector k, row-major. */
void oracle_sym3_eig(const double *s, double *lam, double *V) {
    double A[3][3] = {{s[0], s[3], s[5]}, {s[3], s[1], s[4]}, {s[5], s[4], s[2]}};
    double Q[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
    for (int sweep = 0; sweep < 10; ++sweep) {
        double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        double dia = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
        if (off <= 1e-30 * dia || off == 0.0) break;
        for (int p = 0; p < 2; ++p) {
            for (int q = p + 1; q < 3; ++q) {
                double apq = A[p][q];
                if (apq == 0.0) continue;
                double th = (A[q][q] - A[p][p]) / (2.0 * apq);
                double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < 3; ++k) {
                    double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - sn * akq;
                    A[k][q] = sn * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - sn * aqk;
                    A[q][k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    double qkp = Q[k][p], qkq = Q[k][q];
                    Q[k][p] = c * qkp - sn * qkq;
                    Q[k][q] = sn * qkp + c * qkq;
                }
            }
        }
    }
    int o[3] = {0, 1, 2};
    double d[3] = {A[0][0], A[1][1], A[2][2]};
    /* descending order, stable for ties */
    if (d[o[0]] < d[o[1]]) { int t = o[0]; o[0] = o[1]; o[1] = t; }
    if (d[o[1]] < d[o[2]]) { int t = o[1]; o[1] = o[2]; o[2] = t; }
    if (d[o[0]] < d[o[1]]) { int t = o[0]; o[0] = o[1]; o[1] = t; }
    for (int k = 0; k < 3; ++k) {
        lam[k] = d[o[k]];
        for (int r = 0; r < 3; ++r) V[r * 3 + k] = Q[r][o[k]];
    }
}

static inline double dot3(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* D n for the principal-space elastic matrix [[Dt,Dd,Dd],[Dd,Dt,Dd],[Dd,Dd,Dt]]. */
static inline void dmul3(double Dt, double Dd, const double *n, double *out) {
    out[0] = Dt * n[0] + Dd * n[1] + Dd * n[2];
    out[1] = Dd * n[0] + Dt * n[1] + Dd * n[2];
    out[2] = Dd * n[0] + Dd * n[1] + Dt * n[2];
}

/* Principal-space return of p (descending) in place.  Returns 0 elastic,
 * 1 main plane, 2 edge s1 = s2, 3 edge s2 = s3, 4 apex. */
int oracle_mc3_principal(double *p, double Dt, double Dd, double sphi, double cphi,
                         double spsi, double c) {
    const double k2 = 2.0 * c * cphi;
    double f = (p[0] - p[2]) + (p[0] + p[2]) * sphi - k2;
    if (!(f > 0.0)) return 0;
    const double Ga[3] = {1.0 + sphi, 0.0, -(1.0 - sphi)};
    const double Na[3] = {1.0 + spsi, 0.0, -(1.0 - spsi)};
    double DNa[3], DNb[3];
    dmul3(Dt, Dd, Na, DNa);
    const double a11 = dot3(Ga, DNa);
    double q[3];
    const double g = f / a11;
    for (int k = 0; k < 3; ++k) q[k] = p[k] - g * DNa[k];
    if (q[0] >= q[1] && q[1] >= q[2]) {
        p[0] = q[0]; p[1] = q[1]; p[2] = q[2];
        return 1;
    }
    /* the edge whose ordering constraint the main-plane return crosses first */
    const int edge12 = (1.0 - spsi) * p[0] - 2.0 * p[1] + (1.0 + spsi) * p[2] < 0.0;
    double Gb[3], Nb[3], fb;
    if (edge12) { /* s1 = s2: second plane (s2 - s3) */
        Gb[0] = 0.0; Gb[1] = 1.0 + sphi; Gb[2] = -(1.0 - sphi);
        Nb[0] = 0.0; Nb[1] = 1.0 + spsi; Nb[2] = -(1.0 - spsi);
        fb = (p[1] - p[2]) + (p[1] + p[2]) * sphi - k2;
    } else {      /* s2 = s3: second plane (s1 - s2) */
        Gb[0] = 1.0 + sphi; Gb[1] = -(1.0 - sphi); Gb[2] = 0.0;
        Nb[0] = 1.0 + spsi; Nb[1] = -(1.0 - spsi); Nb[2] = 0.0;
        fb = (p[0] - p[1]) + (p[0] + p[1]) * sphi - k2;
    }
    dmul3(Dt, Dd, Nb, DNb);
    const double a12 = dot3(Ga, DNb), a21 = dot3(Gb, DNa), a22 = dot3(Gb, DNb);
    const double det = a11 * a22 - a12 * a21;
    if (det > 0.0) {
        const double ga = (a22 * f - a12 * fb) / det;
        const double gb = (a11 * fb - a21 * f) / det;
        for (int k = 0; k < 3; ++k) q[k] = p[k] - ga * DNa[k] - gb * DNb[k];
        if ((ga >= 0.0 && gb >= 0.0 && q[0] >= q[2]) || !(sphi > 0.0)) {
            p[0] = q[0]; p[1] = q[1]; p[2] = q[2];
            return edge12 ? 2 : 3;
        }
    }
    const double apex = c * cphi / sphi;
    p[0] = p[1] = p[2] = apex;
    return 4;
}

/* Strength parameters in effect at accumulated plastic strain eq, following
 * the linear softening rule of mohr_coulomb.tcc:8-25 (one peak/residual
 * plastic-strain pair for cohesion, friction and dilation; peak at eq = 0):
 *   cohesion c(eq) = max(c + Hp eq, cr)  (the 2D mcP! rule, mohrcoulomb.jl:36);
 *   when Hp < 0, c > cr and phir > 0, the residual strain is
 *   eq_r = (cr - c) / Hp and, with t = min(eq / eq_r, 1), friction goes
 *   linearly from phi to phir and dilation from psi to 0 (no dilation at the
 *   residual state); otherwise phi and psi stay (the reference's default
 *   phir = 0 leaves them constant). */
void oracle_mc_soften(double eq, double phi, double psi, double coh, double Hp, double cr,
                      double phir, double *phi_out, double *psi_out, double *c_out) {
    *c_out = jl_max(coh + Hp * eq, cr);
    *phi_out = phi;
    *psi_out = psi;
    if (Hp < 0.0 && coh > cr && phir > 0.0) {
        const double eqr = (cr - coh) / Hp;
        const double t = eq >= eqr ? 1.0 : eq / eqr;
        *phi_out = phi + (phir - phi) * t;
        *psi_out = psi - psi * t;
    }
}

/* One particle: s[6] in/out ([xx,yy,zz,xy,yz,xz]), eq in/out, ek out.
 * Returns the return type of oracle_mc3_principal. */
int oracle_mc3_point(double *s, double *eq, double *ek, double Ks, double Gs, double phi0,
                     double psi0, double coh, double Hp, double cr, double phir) {
    double phi, psi, c;
    oracle_mc_soften(*eq, phi0, psi0, coh, Hp, cr, phir, &phi, &psi, &c);
    const double Dt = Ks + 1.333333 * Gs, Dd = Ks - 0.666667 * Gs;
    double lam[3], V[9];
    oracle_sym3_eig(s, lam, V);
    int kind = oracle_mc3_principal(lam, Dt, Dd, sin(phi), cos(phi), sin(psi), c);
    if (kind == 0) {
        *ek = 0.0;
        return 0;
    }
    double sn[6];
    const int ri[6] = {0, 1, 2, 0, 1, 0}, ci[6] = {0, 1, 2, 1, 2, 2};
    for (int k = 0; k < 6; ++k) {
        const int r = ri[k], cc = ci[k];
        sn[k] = lam[0] * V[r * 3 + 0] * V[cc * 3 + 0] + lam[1] * V[r * 3 + 1] * V[cc * 3 + 1] +
                lam[2] * V[r * 3 + 2] * V[cc * 3 + 2];
    }
    double d[6];
    for (int k = 0; k < 6; ++k) {
        d[k] = sn[k] - s[k];
        s[k] = sn[k];
    }
    const double base = 1.0 / ((Dd - Dt) * (2.0 * Dd + Dt));
    const double exx = -(Dd * d[0] + Dt * d[0] - Dd * d[1] - Dd * d[2]) * base;
    const double eyy = -(-Dd * d[0] + Dd * d[1] + Dt * d[1] - Dd * d[2]) * base;
    const double ezz = -(-Dd * d[0] - Dd * d[1] + Dd * d[2] + Dt * d[2]) * base;
    const double exy = d[3] / Gs, eyz = d[4] / Gs, exz = d[5] / Gs;
    *ek = exx + eyy + ezz;
    *eq += sqrt(0.666667 * (exx * exx + eyy * eyy + ezz * ezz +
                            2.0 * (exy * exy + eyz * eyz + exz * exz)));
    return kind;
}

/* ---------------------------------------------------------------------------
 * mcP! 2D plane strain (src/materials/mohrcoulomb.jl:22-86); 3D: the
 * build-defined return above.
 * ------------------------------------------------------------------------- */
void oracle_mcP(oracle_model *m) {
    const int64_t np = m->np;
    if (m->dim == 3) {
        #pragma omp parallel for if (par(m)) schedule(static)
        for (int64_t ix = 0; ix < np; ++ix) {
            int64_t k = m->nid[ix] - 1;
            double s[6];
            for (int j = 0; j < 6; ++j) s[j] = COL(m->sig, np, ix, j);
            double eq = m->eps_q[ix], ek = m->eps_k[ix];
            oracle_mc3_point(s, &eq, &ek, m->Ks[k], m->Gs[k], m->phi[k], m->psi[k], m->c[k],
                             m->Hp[k], m->cr[k], m->phir ? m->phir[k] : 0.0);
            for (int j = 0; j < 6; ++j) COL(m->sig, np, ix, j) = s[j];
            m->eps_q[ix] = eq;
            m->eps_k[ix] = ek;
            deviatoric(m, ix);
        }
        return;
    }
    #pragma omp parallel for if (par(m)) schedule(static)
    for (int64_t ix = 0; ix < np; ++ix) {
        int64_t k = m->nid[ix] - 1;
        double c = m->c[k], Hp = m->Hp[k], cr = m->cr[k], phi = m->phi[k];
        double Gs = m->Gs[k], Ks = m->Ks[k];
        double *s = m->sig;
        c = jl_max(c + Hp * m->eps_q[ix], cr);
        double ds = COL(s, np, ix, 0) - COL(s, np, ix, 1);
        double tau = sqrt(0.25 * (ds * ds) + COL(s, np, ix, 3) * COL(s, np, ix, 3));
        double sg = 0.5 * (COL(s, np, ix, 0) + COL(s, np, ix, 1));
        double f = tau + sg * sin(phi) - c * cos(phi);
        double sn1 = COL(s, np, ix, 0), sn2 = COL(s, np, ix, 1);
        double sn3 = COL(s, np, ix, 2), sn4 = COL(s, np, ix, 3);
        double beta = fabs(c * cos(phi) - sg * sin(phi)) / tau;
        double dsigA = 0.5 * beta * ds;
        double dsigB = c / tan(phi);
        if ((sg <= dsigB) && (f > 0.0)) {
            sn1 = sg + dsigA;
            sn2 = sg - dsigA;
            sn4 = beta * COL(s, np, ix, 3);
        }
        if ((sg > dsigB) && (f > 0.0)) {
            sn1 = dsigB;
            sn2 = dsigB;
            sn4 = 0.0;
        }
        double d1 = sn1 - COL(s, np, ix, 0), d2 = sn2 - COL(s, np, ix, 1);
        double d3 = sn3 - COL(s, np, ix, 2), d4 = sn4 - COL(s, np, ix, 3);
        COL(s, np, ix, 0) = sn1; COL(s, np, ix, 1) = sn2;
        COL(s, np, ix, 2) = sn3; COL(s, np, ix, 3) = sn4;
        double Dt = Ks + 1.333333 * Gs, Dd = Ks - 0.666667 * Gs;
        double base = 1.0 / ((Dd - Dt) * (2.0 * Dd + Dt));
        double exx = -(Dd * d1 + Dt * d1 - Dd * d2 - Dd * d3) * base;
        double eyy = -(-Dd * d1 + Dd * d2 + Dt * d2 - Dd * d3) * base;
        double ezz = -(-Dd * d1 - Dd * d2 + Dd * d3 + Dt * d3) * base;
        double exy = d4 / Gs;
        m->eps_k[ix] = exx + eyy + ezz;
        m->eps_q[ix] += sqrt(0.666667 * (exx * exx + eyy * eyy + ezz * ezz + exy * exy * 2.0));
        deviatoric(m, ix);
    }
}

/* ---------------------------------------------------------------------------
 * vollock1_OS! / vollock2_OS! (utils_OS.jl:1162-1256)
 * ------------------------------------------------------------------------- */
void oracle_vollock(oracle_model *m) {
    const int64_t np = m->np;
    #pragma omp parallel for if (par(m)) schedule(static)
    for (int64_t ix = 0; ix < np; ++ix) {
        int64_t p = m->p2c[ix] - 1;
        double vol = m->omega[ix];
        scatter_add(m, &m->g_sigm[p], vol * m->sigm[ix]);
        scatter_add(m, &m->g_omega[p], vol);
    }
    #pragma omp parallel for if (par(m)) schedule(static)
    for (int64_t ix = 0; ix < np; ++ix) {
        int64_t p = m->p2c[ix] - 1;
        double sm = m->g_sigm[p] / m->g_omega[p];
        COL(m->sig, np, ix, 0) = COL(m->sij, np, ix, 0) + sm;
        COL(m->sig, np, ix, 1) = COL(m->sij, np, ix, 1) + sm;
        COL(m->sig, np, ix, 2) = COL(m->sij, np, ix, 2) + sm;
        COL(m->sig, np, ix, 3) = COL(m->sij, np, ix, 3);
        if (m->dim == 3) {
            COL(m->sig, np, ix, 4) = COL(m->sij, np, ix, 4);
            COL(m->sig, np, ix, 5) = COL(m->sij, np, ix, 5);
        }
        deviatoric(m, ix);
    }
}

/* G2P_OS! and the constitutive update of one step (the same block in
 * OS_MUSL.jl:85-96, OS_USL.jl:123-138,346-361 and OS_USF.jl:25-40,68-83). */
static void g2p_and_stress(oracle_model *m, double dt, double Ti) {
    oracle_g2p(m, dt);
    if (m->constitutive == 0) {
        oracle_liE(m);
    } else if (m->constitutive == 1) {
        oracle_liE(m);
        if (Ti >= m->Te) oracle_dpP(m);
    } else if (m->constitutive == 2) {
        oracle_liE(m);
        if (Ti >= m->Te) oracle_mcP(m);
    }
}

/* ---------------------------------------------------------------------------
 * procedure!(..., Val(:OS), Val(:MUSL), Val(:a|:v)) (src/solvers/OS_MUSL.jl:12-102)
 * procedure!(..., Val(:OS), Val(:USL), Val(:a|:v))  (src/solvers/OS_USL.jl:102-144,325-366)
 * procedure!(..., Val(:OS), Val(:USF), Val(:a|:v))  (src/solvers/OS_USF.jl:11-54,56-97)
 * USF updates the stress FIRST, with the nodal displacement and the basis of
 * the previous step (zero displacement and the initial basis at the first
 * step), and runs MVL after the position update.
 * ------------------------------------------------------------------------- */
void oracle_procedure(oracle_model *m, double dt, double Ti) {
#ifdef _OPENMP
    if (m->nthreads > 1) omp_set_num_threads(m->nthreads);
#endif
    double G = Ti < m->Te ? m->gravity / m->Te * Ti : m->gravity;
    m->Ti_now = Ti;
    if (m->scheme == 2) g2p_and_stress(m, dt, Ti);
    oracle_resetgrid(m);
    oracle_resetmp(m);
    oracle_p2g(m, G);
    oracle_solvegrid(m, dt);
    oracle_dm1(m, dt);
    if (m->scheme == 0) {
        oracle_dm2(m);
        oracle_dm3(m, dt);
    }
    if (m->scheme != 2) g2p_and_stress(m, dt, Ti);
    if (m->mvl) oracle_vollock(m);
}

/* Run the submit_work! loop body nsteps times starting at Ti, accumulating
 * Ti += dt in FP64 exactly like src/solver.jl:118-124.  Returns the final Ti. */
double oracle_run(oracle_model *m, double dt, double Ti, int64_t nsteps) {
    for (int64_t s = 0; s < nsteps; ++s) {
        oracle_procedure(m, dt, Ti);
        Ti += dt;
    }
    return Ti;
}

int oracle_struct_size(void) { return (int)sizeof(oracle_model); }
