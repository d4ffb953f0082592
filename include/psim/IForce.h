/* PSim::IForce -- the force plugin contract
 * (ClusteredCore/include/interfaces/IForce.h:13-65).
 *
 * Same virtual functions in the same order (destructor, getAcceleration,
 * isTimeDependent, quench, then postRoutine under WITHPOST) and the same
 * protected std::string name, so a host compiled against the reference's
 * header calls a GPU plugin through an identical vtable.  A plugin .so
 * exports  extern "C" PSim::IForce *getForce(PSim::config *). */
#ifndef PSIM_COMPAT_IFORCE_H
#define PSIM_COMPAT_IFORCE_H
#include <string>
#include <tuple>
#include <vector>

#include "config.h"
#include "systemState.h"

namespace PSim {
class IForce {
protected:
    std::string name;

public:
    static const int version = 1;
    virtual ~IForce() {}
    /* force on sorted particle `index`, written to particleForce[3*orig..+2]
     * where orig = get<1>((*particleHashIndex)[index]) */
    virtual void getAcceleration(int index, double *sortedParticles, double *particleForce,
                                 std::vector<std::tuple<int, int>> *particleHashIndex,
                                 std::vector<std::tuple<int, int>> *cellStartEnd,
                                 systemState *state) = 0;
    virtual bool isTimeDependent() = 0;
    virtual void quench(systemState *state) { (void)state; }
    std::string getName() { return name; }
#ifdef WITHPOST
    virtual void postRoutine(int index, double *sortedParticles, double *particleForce,
                             std::vector<std::tuple<int, int>> *particleHashIndex,
                             std::vector<std::tuple<int, int>> *cellStartEnd,
                             systemState *state) = 0;
#endif
};
typedef IForce *create_Force(config *);
}  // namespace PSim
#endif
