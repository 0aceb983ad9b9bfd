// Host post-processing of docking results (see ud_post.cpp).
#pragma once
#include <vector>

namespace ud {

static constexpr float kEpsFloat = 1.1920928955078125e-07f;  // epsilon_fl (common.h:227-230)

float rmsd_upper_bound(const float* a, const float* b, int n);
float rmsd_lower_bound(const float* a, const float* b, const int* element, int n);
void cluster_modes(int n_cand, int n_heavy, const float* xyz, const float* energy, float min_rmsd, int max_size,
                   std::vector<int>& kept);
float conf_independent(float x, float num_tors, float rot_internal);
int xs_element(int xs_type);
int ad4_element(int ad_type);

}  // namespace ud
