/* PSim::systemState -- the POD shared by the host and the plugins
 * (ClusteredCore/include/structs/systemState.h:7-19): same members, same
 * order, so a host-built object is read correctly by a GPU plugin. */
#ifndef PSIM_COMPAT_SYSTEM_STATE_H
#define PSIM_COMPAT_SYSTEM_STATE_H
namespace PSim {
struct systemState {
    int nParticles;
    double concentration;
    int boxSize;
    int cellSize;
    int cellScale;
    double temp;
    double currentTime;
    double dTime;
    int seed;
    int outputFreq;
    double endTime;
};
}  // namespace PSim
#endif
