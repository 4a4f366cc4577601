// Variant choices of the iteration passes per grid size, written by
// s_spi_spring_amd/autotune.py from spr_autotune timings on an MI355X (gfx950);
// sizes not listed use variant 0.
#pragma once

namespace spr {

struct TunedVariants {
    int n, col, row;
};
constexpr TunedVariants kTuned[] = {
    {256, 1, 3},
    {384, 1, 3},
    {512, 6, 3},
    {768, 4, 4},
    {1024, 4, 6},
    {2048, 2, 5},
};

}  // namespace spr
