/*
 * cg.c -- oracle restatement of CGMethod::minimize with the backtracking and quadratic line
 * searches, and of the MController wiring with the recovery move. TEST INFRASTRUCTURE ONLY (see
 * oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* CGMethod.hpp:98-107 (double build: LAMBDATOL fixed at 1e-8) */
#define LAMBDAREDUCE 0.5
#define LAMBDATOL 1e-8
#define BACKTRACKSLOPE 0.4

/* std::max / std::min semantics */
static inline double rmax(double a, double b) { return (a < b) ? b : a; }
static inline double rmin(double a, double b) { return (b < a) ? b : a; }

typedef struct {
    int ndof, nvar;
    double *coordLS, *coordMinE, *force, *forcePrev, *searchDir;
    int have_min_e;
    double minimumE, totalEnergy, maxForceBackup;
} cg_state;

static double dot_n(const double* a, const double* b, int n) {
    double r = 0.0;
    for (int i = 0; i < n; i++) r += a[i] * b[i];
    return r;
}

/* maxF (CGMethod.hpp:202-212): component-wise max |F| over numDof; non-finite is an error */
static int max_f(const cg_state* st, double* out) {
    double m = 0;
    for (int i = 0; i < st->ndof; i++) m = rmax(m, fabs(st->force[i]));
    if (!isfinite(m)) return MDN_ERR_NONFINITE;
    *out = m;
    return MDN_OK;
}

static int below_relaxed_tol(const cg_state* st, double factor, double tol) {
    for (int i = 0; i < st->ndof; i++)
        if (fabs(st->force[i]) > factor * tol) return 0;
    return 1;
}

static void shift_search_dir(cg_state* st, double d) {
    /* CGMethod.cpp:356-360 */
    for (int i = 0; i < st->ndof; i++) st->searchDir[i] = st->force[i] + d * st->searchDir[i];
}

typedef struct { double lambda; int numEnergyCall; double newEnergy; } ls_result;

/* backtrackingLineSearch (CGMethod.hpp:238-373); the running-average manager is a fresh object
 * per call with probability 0, so it never changes lambda (LineSearch.hpp:21-48). */
static ls_result backtracking(cg_state* st, const double* coord, const or_cg_callbacks* cb,
                              double MAXDIST, double maxForce, double LAMBDAMAX, double armijo) {
    ls_result res = {0.0, 0, 0.0};
    if (maxForce == 0.0) return res;
    double lambda = rmin(LAMBDAMAX, MAXDIST / maxForce);
    const double currentEnergy = cb->energy(cb->user, coord);
    ++res.numEnergyCall;
    const double expectedSlope = dot_n(st->searchDir, st->force, st->ndof);
    const double reducedSlope = armijo * expectedSlope;
    int conv = 0;
    double energyChange = 0.0, energyLambda = 0.0;
    while (!conv) {
        for (int i = 0; i < st->ndof; i++) st->coordLS[i] = coord[i] + lambda * st->searchDir[i];
        energyLambda = cb->energy(cb->user, st->coordLS);
        ++res.numEnergyCall;
        const double idealEnergyChange = -reducedSlope * lambda;
        if (lambda < 0 || idealEnergyChange > 0) {
            conv = 1;
            lambda = 0.0;
        }
        energyChange = energyLambda - currentEnergy;
        if (energyChange <= idealEnergyChange) conv = 1;
        else lambda *= LAMBDAREDUCE;
        if (lambda <= 0.0 || lambda <= LAMBDATOL) {
            conv = 1;
            lambda = 0.0;
            energyLambda = currentEnergy;
            energyChange = 0;
        }
    }
    (void)energyChange;
    st->totalEnergy = energyLambda;
    res.lambda = lambda;
    res.newEnergy = energyLambda;
    return res;
}

/* quadraticLineSearchV2 (CGMethod.hpp:522-656) and quadraticoptimization (:660-764), with every
 * trial evaluated at coord + lambda * searchDir (the reference moves coord in place and evaluates
 * coordLineSearch, which it does not refresh). Two guards the reference lacks, both on paths where
 * it would evaluate a non-finite or endless trial: a non-finite parabola vertex returns the middle
 * point x1 without an evaluation (the reference evaluates it, then returns x1 through its "out of
 * range" branch), and the parabola iterations stop at QUAD_ITER_CAP, returning x1. */
#define QUADTOL_LAMBDA 1e-3   /* LAMBDAQUADTOL (CGMethod.hpp:106) */
#define QUAD_ITER_CAP 64

static double quad_opt(cg_state* st, const double* coord, const or_cg_callbacks* cb, int* numEnergyCall,
                       double x0, double x1, double x2, double y0, double y1, double y2) {
    const double btlambda = x1;
    for (int iter = 1;; iter++) {
        if (iter > QUAD_ITER_CAP) return x1;
        const double d = (x0 - x1) * (x0 - x2) * (x1 - x2);
        const double a = y0 * (x1 - x2) - y1 * (x0 - x2) + y2 * (x0 - x1);
        const double b = -(y0 * (x1 + x2) * (x1 - x2) - y1 * (x0 + x2) * (x0 - x2) + y2 * (x0 + x1) * (x0 - x1));
        const double lq = -b / (2 * a);
        if (!isfinite(lq)) return x1;
        if (fabs(lq - x1) / x1 < QUADTOL_LAMBDA) return lq;
        if (a / d < 0) return btlambda;   /* the model has no minimum */
        for (int i = 0; i < st->ndof; i++) st->coordLS[i] = coord[i] + lq * st->searchDir[i];
        const double eq = cb->energy(cb->user, st->coordLS);
        ++*numEnergyCall;
        if (fabs(eq - y1) / y1 < QUADTOL_LAMBDA) return lq;
        if (lq < x1) {
            if (eq < y0 && eq < y1) { x1 = lq; y1 = eq; }
            else { x0 = lq; y0 = eq; }
        } else if (x1 < lq) {
            if (eq < y1 && eq < y2) { x1 = lq; y1 = eq; }
            else { x2 = lq; y2 = eq; }
        } else {
            return x1;
        }
    }
}

static ls_result quadratic(cg_state* st, const double* coord, const or_cg_callbacks* cb, double MAXDIST,
                           double maxForce, double LAMBDAMAX) {
    ls_result res = {0.0, 0, 0.0};
    if (maxForce == 0.0) return res;
    double lambda = rmin(LAMBDAMAX, MAXDIST / maxForce);
    const double currentEnergy = cb->energy(cb->user, coord);
    ++res.numEnergyCall;
    const double slope = dot_n(st->searchDir, st->force, st->ndof);
    double energyLambda = 0.0, prevEnergy = 0.0, prevLambda = 0.0;
    /* the recorded (lambda, energy) list: its first entry and the largest energy of the entries
     * after it, which is all the quadratic phase reads of it */
    double lv0 = 0.0, ev0 = 0.0, emax = -INFINITY;
    int n = 0, iter = 0, conv = 0;
    while (!conv) {
        iter++;
        for (int i = 0; i < st->ndof; i++) st->coordLS[i] = coord[i] + lambda * st->searchDir[i];
        energyLambda = cb->energy(cb->user, st->coordLS);
        ++res.numEnergyCall;
        const double idealEnergyChange = -BACKTRACKSLOPE * lambda * slope;
        if (idealEnergyChange > 0) {   /* the reference exits the program; a failed search here */
            lambda = 0.0;
            break;
        }
        if (energyLambda - currentEnergy <= idealEnergyChange) conv = 1;
        else {
            prevLambda = lambda;
            prevEnergy = energyLambda;
            lambda *= LAMBDAREDUCE;
        }
        double l, e;
        if (lambda <= 0.0 || lambda <= LAMBDATOL) {
            conv = 1;
            lambda = 0.0;
            l = lambda; e = currentEnergy;
        } else if (conv) {
            l = lambda; e = energyLambda;
        } else {
            l = prevLambda; e = prevEnergy;
        }
        if (n++ == 0) { lv0 = l; ev0 = e; }
        else emax = rmax(emax, e);
    }
    /* quadstatus: some entry after the first holds more energy than the accepted trial; the
     * reference's loop leaves its index at 0, so the third point is the first entry (the first
     * trial) */
    if (lambda > 0.0 && iter > 1 && energyLambda < currentEnergy && energyLambda < emax) {
        const double lq = quad_opt(st, coord, cb, &res.numEnergyCall, 0.0, lambda, lv0, currentEnergy, energyLambda, ev0);
        lambda = (lq <= 0.0 || lq <= LAMBDATOL) ? 0.0 : lq;
    }
    st->totalEnergy = lambda > 0 ? energyLambda : currentEnergy;
    res.lambda = lambda;
    res.newEnergy = energyLambda;
    return res;
}

static int call_limit_exceeded(const mdn_min_result* r) {
    return (r->energy_call_limit != 0 && r->num_energy_call > r->energy_call_limit) ||
           (r->force_call_limit != 0 && r->num_force_call > r->force_call_limit);
}
static int converged(const mdn_min_result* r) {
    return r->grad_inf_norm < r->grad_tol || r->energy_rel_change < r->energy_rel_tol;
}

int or_cg_minimize(const mdn_cg_params* p, double* coord, int ndof, int nvar,
                   const or_cg_callbacks* cb, double* force_out, mdn_min_result* res) {
    memset(res, 0, sizeof(*res));
    res->energy_call_limit = p->energy_call_limit;
    res->grad_tol = p->grad_tol;
    res->energy_rel_tol = p->energy_rel_tol;
    {
        int beadMaxStep = 3 * cb->n_beads;
        res->force_call_limit = p->force_call_limit > beadMaxStep ? p->force_call_limit : beadMaxStep;
    }
    res->grad_inf_norm = INFINITY;
    res->energy_rel_change = INFINITY;
    const int method = p->method;

    cg_state st;
    memset(&st, 0, sizeof(st));
    st.ndof = ndof;
    st.nvar = nvar;
    st.forcePrev = calloc(ndof > 0 ? ndof : 1, sizeof(double));
    st.searchDir = calloc(ndof > 0 ? ndof : 1, sizeof(double));
    st.coordLS = calloc(nvar > 0 ? nvar : 1, sizeof(double));
    st.force = calloc(nvar > 0 ? nvar : 1, sizeof(double));
    st.coordMinE = calloc(nvar > 0 ? nvar : 1, sizeof(double));
    st.minimumE = INFINITY;  /* startMinimization (CGMethod.cpp:362-366) */
    int status = MDN_OK;

    cb->energy_individual(cb->user, coord, &res->energies_before);
    ++res->num_energy_call;
    cb->force(cb->user, coord, st.force, nvar);
    ++res->num_force_call;
    if ((status = max_f(&st, &res->grad_inf_norm)) != MDN_OK) goto done;
    double prevEnergy = res->energies_before.total;

    int safeMode = 0, numIter = 0;
    double lastFDotF = 0;
#define RESET_CG()                                                                 \
    do {                                                                           \
        shift_search_dir(&st, 0);                                                  \
        if (method == MDN_CG_FLETCHER_REEVES || method == MDN_CG_POLAK_RIBIERE)    \
            lastFDotF = dot_n(st.force, st.force, st.nvar);                        \
        if (method == MDN_CG_POLAK_RIBIERE)                                        \
            memcpy(st.forcePrev, st.force, sizeof(double) * st.ndof);              \
    } while (0)
    RESET_CG();

    while (!call_limit_exceeded(res) && !converged(res)) {
        double beta = 0, newGrad = 0, prevGrad = 0;
        numIter++;
        /* CGMethod.hpp:1319-1331: QUADRATIC backtracks with slope 0 in safe mode */
        ls_result lsr = (p->line_search == MDN_LS_QUADRATIC && !safeMode)
                            ? quadratic(&st, coord, cb, p->max_dist, res->grad_inf_norm, p->lambda_max)
                            : backtracking(&st, coord, cb, p->max_dist, res->grad_inf_norm, p->lambda_max,
                                           safeMode ? 0 : BACKTRACKSLOPE);
        const double lambda = lsr.lambda;
        res->num_energy_call += lsr.numEnergyCall;
        if (lambda == 0) {
            if (p->try_recover) {
                if (safeMode) {
                    if (cb->recovery) {
                        status = cb->recovery(cb->user, coord, st.searchDir, ndof);
                        if (status != MDN_OK) goto done;
                    }
                    res->num_recovery++;
                    cb->force(cb->user, coord, st.force, nvar);
                    ++res->num_force_call;
                    if ((status = max_f(&st, &res->grad_inf_norm)) != MDN_OK) goto done;
                    RESET_CG();
                    safeMode = 0;
                } else {
                    RESET_CG();
                    safeMode = 1;
                    res->num_safe_mode++;
                }
            } else {
                res->error_flags |= MDN_MIN_ERR_LINESEARCH;
                res->num_iter = numIter;
                goto done_early;
            }
        } else {
            for (int i = 0; i < ndof; i++) coord[i] += lambda * st.searchDir[i];
            cb->force(cb->user, coord, st.force, nvar);
            ++res->num_force_call;
            if ((status = max_f(&st, &res->grad_inf_norm)) != MDN_OK) goto done;
            res->energy_rel_change = fabs(lsr.newEnergy - prevEnergy) / fabs(prevEnergy);
            prevEnergy = lsr.newEnergy;
            if (method == MDN_CG_FLETCHER_REEVES) {
                newGrad = dot_n(st.force, st.force, nvar);
                beta = newGrad / lastFDotF;
            } else if (method == MDN_CG_POLAK_RIBIERE) {
                newGrad = dot_n(st.force, st.force, nvar);
                prevGrad = dot_n(st.force, st.forcePrev, ndof);
                beta = rmax(0.0, (newGrad - prevGrad) / lastFDotF);
            }
            shift_search_dir(&st, beta);
            if (dot_n(st.searchDir, st.force, ndof) <= 0) RESET_CG();
            if (method == MDN_CG_POLAK_RIBIERE) memcpy(st.forcePrev, st.force, sizeof(double) * ndof);
            if (below_relaxed_tol(&st, 10, p->grad_tol) && numIter > res->force_call_limit / 2) {
                /* copycoordsifminimumE (CGMethod.hpp:794-803) */
                if (st.totalEnergy <= st.minimumE) {
                    st.minimumE = st.totalEnergy;
                    st.maxForceBackup = res->grad_inf_norm;
                    memcpy(st.coordMinE, coord, sizeof(double) * nvar);
                    st.have_min_e = 1;
                }
            }
            if (method == MDN_CG_FLETCHER_REEVES || method == MDN_CG_POLAK_RIBIERE) lastFDotF = newGrad;
        }
    }
#undef RESET_CG
    res->num_iter = numIter;
    res->call_limit_exceeded = call_limit_exceeded(res);
    if (p->report_on_call_limit && res->call_limit_exceeded) {
        if (st.have_min_e) {  /* copybackupcoordinates (CGMethod.hpp:805-814) */
            memcpy(coord, st.coordMinE, sizeof(double) * nvar);
            res->restored_min_energy = 1;
        }
    }
    cb->energy_individual(cb->user, coord, &res->energies_after);
    ++res->num_energy_call;
done_early:
    res->converged_grad = res->grad_inf_norm < res->grad_tol;
    res->converged_energy = res->energy_rel_change < res->energy_rel_tol;
    res->call_limit_exceeded = call_limit_exceeded(res);
    if (force_out) memcpy(force_out, st.force, sizeof(double) * nvar);
done:
    free(st.forcePrev); free(st.searchDir); free(st.coordLS); free(st.force); free(st.coordMinE);
    return status;
}

/* ---- MController wiring ----------------------------------------------------------- */

typedef struct { or_system* s; double max_dist; } sys_user;

static double sys_energy(void* u, const double* coord) {
    return or_ffm_energy(((sys_user*)u)->s, coord, NULL);
}
static void sys_force(void* u, const double* coord, double* f, int nvar) {
    (void)nvar;
    or_ffm_forces(((sys_user*)u)->s, coord, f);
}
static void sys_energy_ind(void* u, const double* coord, mdn_energy_report* rep) {
    or_ffm_energy(((sys_user*)u)->s, coord, rep);
}
/* moveAlongDirectionLimitingBeadMove_ (MController.h:273-339) */
static int sys_recovery(void* u, double* coord, const double* dir, int ndof) {
    double step = INFINITY;
    double maxd = ((sys_user*)u)->max_dist;
    for (int i = 0; i < ndof; i++) {
        double v = maxd / fabs(dir[i]);
        if (v < step) step = v;
    }
    if (step == INFINITY) return MDN_ERR_RECOVERY;
    for (int i = 0; i < ndof; i++) coord[i] += step * dir[i];
    return MDN_OK;
}

int or_system_minimize(or_system* s, const mdn_cg_params* p, double* coord, double* force_out,
                       mdn_min_result* res) {
    sys_user u = {s, p->max_dist};
    or_cg_callbacks cb = {sys_energy, sys_force, sys_energy_ind, sys_recovery, &u, s->n_beads};
    int ndof = 3 * s->n_beads;
    int st = or_cg_minimize(p, coord, ndof, ndof, &cb, force_out, res);
    if (st != MDN_OK) return st;
    /* MController.h:142-155: a line-search error runs a diagnosis and throws */
    if (res->error_flags & MDN_MIN_ERR_LINESEARCH) return MDN_ERR_LINESEARCH;
    return MDN_OK;
}
