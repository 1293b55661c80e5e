// Library entry point of the DTFE drop-in: the declaration of the
// reference's src/DTFE.h:36-40, served by the MI355X kernels.
#ifndef DTFE_HEADER
#define DTFE_HEADER

#include <stdexcept>
#include <string>
#include <vector>

#include "define.h"
#include "particle_data.h"
#include "quantities.h"
#include "user_options.h"

// Computes the grid interpolation. Clears *allParticles, as the reference does.
void DTFE(std::vector<Particle_data> *allParticles, std::vector<Sample_point> &samples, User_options &userOptions,
          Quantities *uQuantities, Quantities *aQuantities);

namespace dtfe_dropin {
// The reference reports fatal errors and exits (src/message.h:57-65); the
// drop-in does the same by default. With throw_on_error(true) it throws
// dtfe_dropin::Error instead, for callers that need to recover.
struct Error : std::runtime_error {
    explicit Error(const std::string &m) : std::runtime_error(m) {}
};
void throw_on_error(bool enable);
// Seconds spent in the host tessellation by the last DTFE() call.
double last_tessellation_seconds();
}  // namespace dtfe_dropin

#endif
