// Compiles the C++20 facade (cppwrf_amd/csrc/dyn_facade.hpp) the way a reference caller would include it, links
// the library, and exercises what runs without a GPU: the sound-step rule, parameter validation, and the error
// path of a context on a machine with no device (std::runtime_error with the library's message).
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../cppwrf_amd/csrc/dyn_facade.hpp"

int main() {
  namespace dyn = cppwrf_amd::dyn;
  // reference RK3::auto_sound_steps: max(2*(int(300*dt/min(dx,dy) - 0.01) + 1), 4)
  if (dyn::RK3::auto_sound_steps(6.0, 1000.0, 1000.0) != 4) return 2;
  if (dyn::RK3::auto_sound_steps(90.0, 15000.0, 15000.0) != 4) return 3;
  if (dyn::RK3::auto_sound_steps(60.0, 3000.0, 3000.0) != 12) return 4;
  try {
    dyn::Context ctx(0);
    // a GPU is present: a domain with an unsupported advection order must be refused
    mi_grid g{16, 16, 10, MI_BC_PERIODIC, MI_BC_PERIODIC, 1000.0, 1000.0, 6.0};
    dyn::LargeStepParams lp;
    lp.h_adv_order = 4;
    dyn::AcousticParams ap;
    mi_levels lv{};
    try {
      dyn::RK3 rk(ctx, g, lp, ap, 4, lv);
    } catch (const std::runtime_error& e) {
      std::printf("refused: %s\n", e.what());
      return 0;
    }
    return 5;
  } catch (const std::runtime_error& e) {
    std::printf("no device: %s\n", e.what());
    return std::string(e.what()).find("cppwrf_amd") == 0 ? 0 : 6;
  }
}
