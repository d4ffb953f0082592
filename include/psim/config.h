/* PSim::config as seen by a plugin (ClusteredCore/include/config.h:15-63).
 *
 * The host owns the object and hands its address to getForce(config*); a
 * plugin reads parameters through getParam<T>(key, default), whose
 * specialisations each plugin compiles itself (as the reference's plugins
 * carry their own copy of config.cpp:97-175).  The data members are declared
 * in the reference's order (options, then suppressOutput) so the host's
 * object is read in place. */
#ifndef PSIM_COMPAT_CONFIG_H
#define PSIM_COMPAT_CONFIG_H
#include <map>
#include <string>

namespace PSim {
class config {
    std::map<std::string, std::string> options;
    bool suppressOutput;

public:
    static const int version = 1;
    /* plugins never construct a config; these exist so host-side test
     * programs can build one */
    config() : suppressOutput(true) {}
    void set(const std::string &key, const std::string &value) { options[key] = value; }
    bool containsKey(const std::string &key) { return options.count(key) != 0; }
    void showOutput() { suppressOutput = false; }
    void hideOutput() { suppressOutput = true; }
    /* typed lookup with default; int/double/float use std::stoi/stod/stof,
     * which read the leading number (config.cpp:97-155) */
    template <typename T>
    T getParam(std::string key, T def);
};
}  // namespace PSim
#endif
