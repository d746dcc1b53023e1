// PERIODIC mode, fused path (placeholder until the tiled kernel lands).
#include "mpd_grid.hpp"

namespace mpd {

bool periodic_fused_supported(const mpd_grid_s*) { return false; }

int periodic_fused_step(mpd_grid_s*)
{
    return fail(MPD_ERR_UNSUPPORTED, "fused periodic path not available");
}

}  // namespace mpd
