/* PSim::type3 -- three-component value (ClusteredCore/include/structs/type3.h:6-16).
 * Same member order and types as the reference so objects cross the plugin
 * boundary unchanged. */
#ifndef PSIM_COMPAT_TYPE3_H
#define PSIM_COMPAT_TYPE3_H
namespace PSim {
template <class T>
struct type3 {
    T x, y, z;
    type3() : x(0), y(0), z(0) {}
    type3(T a, T b, T c) : x(a), y(b), z(c) {}
};
}  // namespace PSim
#endif
