// Persistent signing engine, Dilithium2 instantiations.
#include "sign_persistent.hpp"

namespace dil {
DIL_DEFINE_PERSISTENT_LAUNCHER(2)
}  // namespace dil
