// lambrex.cpp -- process-wide device context for the AmrSim facade.
#include "lambrex.h"

#include <cstdlib>
#include <mutex>

namespace {
lbx_ctx* g_ctx = nullptr;
std::mutex g_mu;
}  // namespace

lbx_ctx* lambrex::Context() {
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ctx) {
    const char* env = std::getenv("LBX_DEVICE");
    const int dev = env ? std::atoi(env) : 0;
    if (lbx_ctx_create(dev, &g_ctx) != LBX_OK) {
      g_ctx = nullptr;
      throw lambrex::Error(std::string("lambrexInit: ") + lbx_last_error(nullptr));
    }
  }
  return g_ctx;
}

void lambrexInit() { (void)lambrex::Context(); }

void lambrexFinalise() {
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_ctx) {
    lbx_ctx_destroy(g_ctx);
    g_ctx = nullptr;
  }
}
