/* The CPU oracle and the seeded network generator under AddressSanitizer and UBSan (built by
 * tests/test_oracle_sanitized.py with -fsanitize=address,undefined): the paths the oracle's
 * other tests take, on the 50-filament network, a dense config-3 network (safe mode, recovery)
 * and a parallel pair (the Simpson fallbacks), so that an out-of-bounds access or undefined
 * arithmetic in the checker surfaces as a sanitizer report. Prints "ok <checks>" on success. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/mdn_synth.h"
#include "../../oracle/oracle.h"

static int run_network(int config, int index, int dense, int ncalls, int* checks) {
    mdn_synth_params sp;
    if (mdn_synth_preset(config, index, dense, &sp) != MDN_OK) return 1;
    mdn_synth* net = NULL;
    if (mdn_synth_create(&sp, &net) != MDN_OK) return 1;
    const mdn_system_desc* d = mdn_synth_desc(net);
    const int nv = 3 * d->n_beads;
    double* x = malloc(sizeof(double) * nv);
    double* f = malloc(sizeof(double) * nv);
    memcpy(x, mdn_synth_coord(net), sizeof(double) * nv);
    int st = 0;
    or_system* s = or_system_create(d, &st);
    if (!s || st) return 1;
    for (int cells = 0; cells < 2; cells++) {   /* brute force and cell list: the same pairs */
        or_system_set_cells(s, cells);
        if (or_system_nl(s, x) != MDN_OK) return 1;
    }
    mdn_energy_report rep;
    const double e = or_ffm_energy(s, x, &rep);
    or_ffm_forces(s, x, f);
    for (int ff = 0; ff < 6; ff++) {
        memset(f, 0, sizeof(double) * nv);
        or_ff_forces(s, ff, x, f);
    }
    mdn_cg_params p;
    memset(&p, 0, sizeof(p));
    p.method = MDN_CG_POLAK_RIBIERE;
    p.try_recover = 1;
    p.max_dist = 1.0;
    p.lambda_max = 0.01;
    p.grad_tol = 1.0;
    p.energy_rel_tol = 1e-9;
    p.force_call_limit = 10000;
    p.energy_call_limit = ncalls;
    p.report_on_call_limit = 1;
    mdn_min_result res;
    st = or_system_minimize(s, &p, x, f, &res);
    if (st != MDN_OK) return 1;
    const double mon[1] = {2.7};
    const int nmon[1] = {40};
    double* lp = calloc((size_t)d->n_filaments * 40 + 1, sizeof(double));
    double* lm = calloc((size_t)d->n_filaments * 40 + 1, sizeof(double));
    int nt = 1;
    for (int q = 0; q < d->n_filaments; q++)
        if (d->fil_type[q] + 1 > nt) nt = d->fil_type[q] + 1;
    if (nt == 1 && or_system_load_forces(s, x, 1, mon, nmon, 40, lp, lm) != MDN_OK) return 1;
    *checks += 1;
    if (!isfinite(e) && e != INFINITY) return 1;
    free(lp); free(lm);
    or_system_destroy(s);
    free(x); free(f);
    mdn_synth_destroy(net);
    return 0;
}

int main(void) {
    int checks = 0;
    if (run_network(0, 0, 0, 40, &checks)) { printf("fail config 0\n"); return 1; }
    if (run_network(3, 4, 1, 30, &checks)) { printf("fail dense config 3\n"); return 1; }
    /* Simpson fallbacks on a parallel coplanar pair (energyN) and a collinear one (forceN) */
    const double par[12] = {0, 0, 0, 100, 0, 0, 0, 30, 0, 100, 30, 0};
    const double col[12] = {0, 0, 0, 60, 0, 0, 80, 0, 0, 140, 0, 0};
    const int bs[4] = {0, 3, 6, 9};
    const double krep[1] = {1e5}, eq[2] = {108.0, 108.0};
    double fo[12];
    const double en = or_exvol_energyN(par, bs, krep, eq, 0, 0);
    memset(fo, 0, sizeof(fo));
    or_exvol_forceN(col, fo, bs, krep, eq, 0, 0);
    if (!(en > 0) || !isfinite(fo[0])) { printf("fail simpson\n"); return 1; }
    checks += 1;
    printf("ok %d\n", checks);
    return 0;
}
