// C++ caller of the drop-in adapters (include/mdn_mech.hpp): the MController-shaped sequence
// serialize/vectorize -> reset neighbor lists -> minimize -> stretch forces and boundary load
// forces, on the seeded 50-filament network. Prints "pairs energy_before energy_after iterations
// max_force converged n_linker_stretch_forces n_loaded_filament_ends".
#include <cmath>
#include <cstdio>
#include <vector>

#include "mdn_mech.hpp"
#include "mdn_synth.h"

int main() {
    mdn_synth_params sp;
    if (mdn_synth_preset(0, 0, 0, &sp) != MDN_OK) return 2;
    mdn_synth* net = nullptr;
    if (mdn_synth_create(&sp, &net) != MDN_OK) return 2;
    const mdn_system_desc* d = mdn_synth_desc(net);
    std::vector<double> coord(mdn_synth_coord(net), mdn_synth_coord(net) + 3 * d->n_beads);
    try {
        mdn::ForceFieldManager ffm(0);
        ffm.vectorizeAllForceFields(*d, coord);
        long pairs = ffm.resetNeighborLists(coord);
        double e0 = ffm.computeEnergy(coord.data());
        std::vector<double> f(coord.size());
        ffm.computeForces(coord.data(), f.data(), (int)f.size());
        mdn::CGMethod cg;
        mdn::ConjugateGradientParams p;
        std::vector<double> force;
        auto res = cg.minimize(ffm, p, coord, &force);
        double mx = 0;
        for (double v : force) mx = std::fmax(mx, std::fabs(v));
        // the post-minimization consumers (MController::afterMinimization_)
        std::vector<double> lsf, msf, lp, lm;
        ffm.assignallforcemags(lsf, msf);
        const int stride = ffm.computeLoadForces(d->n_filaments, {2.7}, {40}, lp, lm);
        int loaded = 0;
        for (int fi = 0; fi < d->n_filaments; fi++) loaded += (lp[(size_t)fi * stride] > 0) + (lm[(size_t)fi * stride] > 0);
        std::printf("%ld %.10e %.10e %d %.6e %d %zu %d\n", pairs, e0, res.energies_after.total, res.num_iter, mx,
                    (int)res.converged(), lsf.size(), loaded);
    } catch (const mdn::Error& e) {
        std::printf("error %d %s\n", e.code, e.what());
        mdn_synth_destroy(net);
        return 3;
    }
    mdn_synth_destroy(net);
    return 0;
}
