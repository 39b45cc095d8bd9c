// Entry points declared in deco_hip.h whose implementation has not landed yet.
#include "common.hpp"

using namespace deco;

#define DECO_STUB(sig) \
  sig { return guard([&] { fail(DECO_ERR_UNSUPPORTED, "%s is not implemented", __func__); }); }

extern "C" {
}
