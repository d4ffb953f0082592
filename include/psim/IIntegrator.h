/* PSim::IIntegrator -- the integrator contract
 * (ClusteredCore/include/interfaces/IIntegrator.h:20-54).
 *
 * Same virtual functions in the same order (destructor, then
 * nextSystem(particle**, systemState*)) and the same protected std::string
 * name, so a host compiled against the reference's header calls a GPU
 * integrator through an identical vtable.  The reference host constructs its
 * integrator directly (ClusteredUI/src/runSim.cpp:127) and hands it to the
 * system (system.cpp:69), which calls nextSystem once per step
 * (system.cpp:196) and deletes it (system.cpp:151).  A GPU integrator
 * library exports  extern "C" PSim::IIntegrator *getIntegrator(PSim::config *)
 * for hosts that load it by name, as they load force plugins. */
#ifndef PSIM_COMPAT_IINTEGRATOR_H
#define PSIM_COMPAT_IINTEGRATOR_H
#include <string>

#include "config.h"
#include "systemState.h"

namespace PSim {
class particle;
class IIntegrator {
protected:
    std::string name;

public:
    static const int version = 1;
    virtual ~IIntegrator() {}
    /* advance every particle one step; 0 = no error */
    virtual int nextSystem(particle **items, systemState *state) = 0;
    std::string getName() { return name; }
};
typedef IIntegrator *create_Integrator(config *);
}  // namespace PSim
#endif
