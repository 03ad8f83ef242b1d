// Persistent signing engine, Dilithium3 instantiations.
#include "sign_persistent.hpp"

namespace dil {
DIL_DEFINE_PERSISTENT_LAUNCHER(3)
}  // namespace dil
