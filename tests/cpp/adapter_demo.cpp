// C++ caller of the drop-in adapters (include/mdn_mech.hpp): the MController-shaped sequence
// serialize/vectorize -> reset neighbor lists -> minimize -> stretch forces and boundary load
// forces, on the seeded 50-filament network, and the per-force-field adapters (ForceField).
// Then the device-resident calls and serializeDof on the device. Prints "pairs energy_before energy_after iterations max_force
// converged n_linker_stretch_forces n_loaded_filament_ends force_field_adapters_ok
// device_resident_ok".
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
        // the per-force-field adapters (ForceField interface): their forces accumulated one after
        // another equal the manager's, their energies sum to its total, in the manager's order
        std::vector<double> facc(coord.size(), 0.0);
        double esum = 0;
        int ff_ok = 1, k = 0;
        for (const auto& ffp : ffm.forceFields()) {
            ffp->vectorize();
            ff_ok &= ffp->getName() == mdn::ff_name(k++);
            esum += ffp->computeEnergy(coord.data());
            ffp->computeForces(coord.data(), facc.data());
        }
        double fdiff = 0, fmax = 0;
        for (size_t i = 0; i < f.size(); i++) {
            fdiff = std::fmax(fdiff, std::fabs(facc[i] - f[i]));
            fmax = std::fmax(fmax, std::fabs(f[i]));
        }
        ff_ok &= fdiff <= 1e-12 * fmax && std::fabs(esum - e0) <= 1e-12 * std::fabs(e0);
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
        auto* lk = dynamic_cast<mdn::GpuLinkerStretching*>(ffm.forceFields()[MDN_FF_LINKER].get());
        lk->assignforcemags();
        ff_ok &= lk->stretchForce() == lsf;
        // device-resident minimization (no coordinate transfer) of a displaced state: the same
        // iterations and bit-identical coordinates as the host-buffer call on the same state
        std::vector<double> xp(coord), xh;
        for (size_t i = 0; i < xp.size(); i++) xp[i] += 2.0 * std::sin(0.7 * (double)i);
        xh = xp;
        ffm.resetNeighborLists(xh);
        auto rh = cg.minimize(ffm, p, xh, nullptr);
        ffm.setCoords(xp);
        ffm.resetNeighborLists();
        auto rd = cg.minimize(ffm, p);
        std::vector<double> xd;
        ffm.getCoords(xd);
        int dev_ok = xd == xh && rd.num_iter == rh.num_iter && rd.converged();
        // seeded bead noise generated on the device: unit variance per coordinate
        ffm.setCoords(xd);
        ffm.perturb(3, 1, 1.0);
        std::vector<double> xn;
        ffm.getCoords(xn);
        double ss = 0;
        for (size_t i = 0; i < xn.size(); i++) ss += (xn[i] - xd[i]) * (xn[i] - xd[i]);
        const double rms = std::sqrt(ss / (double)xn.size());
        dev_ok &= rms > 0.9 && rms < 1.1;
        // serializeDof on the device for the same bead set in reversed order (all beads current),
        // then the next system's upload from those device coordinates
        {
            const int nb = d->n_beads;
            std::vector<int32_t> src(nb), rb(d->fil_bead_offset[d->n_filaments]);
            for (int i = 0; i < nb; i++) src[i] = nb - 1 - i;
            for (size_t k = 0; k < rb.size(); k++) rb[k] = nb - 1 - d->fil_beads[k];
            mdn_system_desc d2 = *d;
            d2.fil_beads = rb.data();
            ffm.serializeDof(src, {});
            ffm.vectorizeAllForceFields(d2);
            std::vector<double> xr;
            ffm.getCoords(xr);
            for (int i = 0; i < nb; i++)
                for (int q = 0; q < 3; q++) dev_ok &= xr[3 * i + q] == xn[3 * (nb - 1 - i) + q];
        }
        std::printf("%ld %.10e %.10e %d %.6e %d %zu %d %d %d\n", pairs, e0, res.energies_after.total, res.num_iter, mx,
                    (int)res.converged(), lsf.size(), loaded, ff_ok, dev_ok);
    } catch (const mdn::Error& e) {
        std::printf("error %d %s\n", e.code, e.what());
        mdn_synth_destroy(net);
        return 3;
    }
    mdn_synth_destroy(net);
    return 0;
}
