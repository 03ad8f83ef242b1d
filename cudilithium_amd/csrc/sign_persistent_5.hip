// Persistent signing engine, Dilithium5 instantiations.
#include "sign_persistent.hpp"

namespace dil {
DIL_DEFINE_PERSISTENT_LAUNCHER(5)
}  // namespace dil
