// Output container with the layout of the reference's Quantities
// (src/quantities.h:56-76): one std::vector per field, empty unless computed.
#ifndef DTFE_DROPIN_QUANTITIES_H
#define DTFE_DROPIN_QUANTITIES_H

#include <vector>

#include "Pvector.h"
#include "define.h"

struct Quantities {
    std::vector<Real> density;
    std::vector<Pvector<Real, noVelComp> > velocity;
    std::vector<Pvector<Real, noGradComp> > velocity_gradient;
    std::vector<Real> velocity_divergence;
    std::vector<Pvector<Real, noShearComp> > velocity_shear;
    std::vector<Pvector<Real, noVortComp> > velocity_vorticity;
    std::vector<Real> velocity_std;
    std::vector<Pvector<Real, noScalarComp> > scalar;
    std::vector<Pvector<Real, noScalarGradComp> > scalar_gradient;

    // size of any non-empty member (0 when all are empty), as src/quantities.cc:108-121
    size_t size() const {
        if (!density.empty()) return density.size();
        if (!velocity.empty()) return velocity.size();
        if (!velocity_gradient.empty()) return velocity_gradient.size();
        if (!velocity_divergence.empty()) return velocity_divergence.size();
        if (!velocity_shear.empty()) return velocity_shear.size();
        if (!velocity_vorticity.empty()) return velocity_vorticity.size();
        if (!velocity_std.empty()) return velocity_std.size();
        if (!scalar.empty()) return scalar.size();
        if (!scalar_gradient.empty()) return scalar_gradient.size();
        return 0;
    }
};

#endif
