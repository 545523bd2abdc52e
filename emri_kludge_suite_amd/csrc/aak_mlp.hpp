// aak_mlp.hpp -- the learned start point of the frequency map (aak_mlp.hip):
// sizes shared with the C ABI and the host stage's inference call.
#pragma once

struct aak_mlp_s;

namespace aak {

constexpr int kMlpHidden = 128;
// W1[4][128] b1[128] W2[128][128] b2[128] W3[128][128] b3[128] W4[128][3] b4[3]
constexpr int kMlpParams = 4 * 128 + 128 + 2 * (128 * 128 + 128) + 3 * 128 + 3;
// feature centres[4], feature scales[4], target means[3], target scales[3]
constexpr int kMlpNorm = 14;

// Start points (v_map, M_map / M, s_map) of the frequency map at n points
// pts = (a, p, e, iota) (row-major n x 4), one batched launch on the model's
// stream; safe to call from several host threads (calls are serialised).
int mlp_infer(aak_mlp_s *m, const double *pts, int n, double *maps);

}  // namespace aak
