// Environment-map tables (host side): the importance-sampling distribution of a latitude-longitude
// environment texture and its selection power, built in FP64 as the reference's EnvironmentLight
// constructor (src/render/light.cpp:61-114) and rounded to float for the device.
#pragma once

#include <cstdint>
#include <vector>

#include "mrg.h"

namespace mrg {

struct EnvTables {
    bool               uniform = true;  // no usable map: uniform sphere sampling, pdf 1 / (4 pi)
    uint32_t           gw = 0, gh = 0;  // grid: the first mip level of at most 1024 x 512 texels
    float              strength = 1;
    float              rotation = 0;   // turns: remainder(degrees, 360) / 360
    float              integral = 0;   // luminance integral of constant + map (light selection power)
    std::vector<float> cdf;            // [gw * gh + 1] texel selection CDF (FP64, rounded)
    std::vector<float> rows;           // [gh][4]: texel solid angle, cos(theta0), cos(theta1), 2 sin((t0+t1)/2) sin((t1-t0)/2)
};

// Tables for `environment` (constant RGB) plus, when `texture` is non-null and strength != 0, the map.
EnvTables build_env_tables(const float environment[3], const mrg_texture* texture, float strength, float rotation_degrees);

}  // namespace mrg
