/* PSim::particle -- the host's per-particle object as an integrator adaptor
 * sees it (ClusteredCore/include/particle.h:14-49).
 *
 * The data members are declared with the reference's types in the
 * reference's order (name, pos, pos0, vel, frc, frc0, r, m, cll,
 * coorNumber, interactions), so a host-built particle is read and written
 * in place.  No member function is declared: the adaptor reaches the
 * members through psim_particle_access, a friend, which leaves the layout
 * unchanged and emits no symbol of PSim::particle. */
#ifndef PSIM_COMPAT_PARTICLE_H
#define PSIM_COMPAT_PARTICLE_H
#include <vector>

#include "type3.h"

struct psim_particle_access;

namespace PSim {
class particle {
    int name;
    type3<double> pos;   /* current position                           */
    type3<double> pos0;  /* previous position, re-imaged next to pos    */
    type3<double> vel;
    type3<double> frc;   /* force of this step (after the push)        */
    type3<double> frc0;  /* force of the previous step                 */
    double r, m;         /* radius and mass                            */
    type3<int> cll;
    int coorNumber;
    std::vector<particle *> interactions;
    friend struct ::psim_particle_access;
};
}  // namespace PSim
#endif
