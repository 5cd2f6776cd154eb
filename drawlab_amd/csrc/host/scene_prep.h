#pragma once
#include <stdint.h>

#include <vector>

namespace dl {

// Normalised area CDF of a mesh (nf + 1 entries); returns 1 / total area.
float area_cdf(const float* V, const uint32_t* F, uint32_t nf, std::vector<float>& cdf);

// Environment importance tables for a lat-long rgb image (row 0 = theta 0).
void env_tables(const float* rgb, int w, int h, std::vector<float>& q, std::vector<uint32_t>& alias,
                std::vector<float>& pdf);

}  // namespace dl
