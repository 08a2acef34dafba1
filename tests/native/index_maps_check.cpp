// CPU check of the spectral index maps in bldfm_types.h (the code the
// kernels run).  For a grid of (nlx, nly, nxe, nye) it verifies that
//  * every shooting lane decodes to an in-range mode, every retained mode
//    (fx, fy) is produced exactly once by a lane or as the mirror of a
//    half-plane lane, and Nyquist indices are unique and < nlx + nly;
//  * the packed-spectrum positions written by the half-plane lanes (and
//    their mirrors), the origin lane and the Nyquist fix-up cover the compact
//    band nxb x nyb exactly once, all in bounds;
//  * compact -> padded maps are injective and inside [0, ne).
// Prints "OK <cases>" and exits 0, or prints the first failure and exits 1.
#include <stdio.h>
#include <stdlib.h>

#include <set>
#include <utility>
#include <vector>

#include "../../bldfm_amd/csrc/bldfm_types.h"

static int fails = 0;
#define CHECK(c, ...)                                  \
  do {                                                 \
    if (!(c)) {                                        \
      printf("FAIL %s:%d: ", __FILE__, __LINE__);      \
      printf(__VA_ARGS__);                             \
      printf("\n");                                    \
      exit(1);                                         \
    }                                                  \
  } while (0)

static void check_case(int nlx, int nly, int nxe, int nye) {
  const int nxb = band_len(nlx, nxe), nyb = band_len(nly, nye);
  CHECK(nxb <= nxe && nyb <= nye, "band longer than padded grid %d %d", nxb, nyb);
  // retained frequencies per axis
  auto in_axis = [](int f, int nl) { return f >= -(nl / 2) && f <= (nl - 1) / 2; };
  std::set<std::pair<int, int>> modes_seen;
  std::vector<int> zcount((size_t)nxb * nyb, 0);
  std::set<int> nyq_seen;
  const int lanes = shoot_lanes(nlx, nly);
  int idle = 0;
  for (int lane = 0; lane < lanes; ++lane) {
    int fx = 0, fy = 0, nidx = -1;
    const int kind = decode_lane(lane, nlx, nly, &fx, &fy, &nidx);
    if (kind == LANE_IDLE) { ++idle; continue; }
    CHECK(in_axis(fx, nlx) && in_axis(fy, nly), "lane %d mode (%d,%d) out of range", lane, fx, fy);
    CHECK(modes_seen.insert({fx, fy}).second, "mode (%d,%d) twice", fx, fy);
    const int cx = compact_of_freq(fx, nxb), cy = compact_of_freq(fy, nyb);
    CHECK(cx >= 0 && cx < nxb && cy >= 0 && cy < nyb, "compact out of range");
    if (kind == LANE_HALF) {
      CHECK(in_axis(-fx, nlx) && in_axis(-fy, nly), "half lane without partner");
      CHECK(modes_seen.insert({-fx, -fy}).second, "mirror (%d,%d) twice", -fx, -fy);
      const int mx = compact_mirror(cx, nxb), my = compact_mirror(cy, nyb);
      CHECK(mx == compact_of_freq(-fx, nxb) && my == compact_of_freq(-fy, nyb), "mirror map");
      zcount[(size_t)cx * nyb + cy]++;
      zcount[(size_t)mx * nyb + my]++;
    } else if (kind == LANE_ORIGIN) {
      CHECK(fx == 0 && fy == 0, "origin");
      zcount[0]++;
    } else {
      CHECK(nidx >= 0 && nidx < nlx + nly, "nyq index %d", nidx);
      CHECK(nyq_seen.insert(nidx).second, "nyq index twice");
      CHECK(nyq_index_at(cx, cy, nlx, nly, nxb, nyb) == nidx, "nyq_index_at mismatch (%d,%d)", fx, fy);
      CHECK(nidx < nyq_count(nlx, nly), "nyq index beyond count");
    }
  }
  CHECK((int)modes_seen.size() == nlx * nly, "modes covered %zu of %d", modes_seen.size(), nlx * nly);
  CHECK((int)nyq_seen.size() == nyq_count(nlx, nly), "nyq count");
  const int fixc = nyq_fix_count(nlx, nly, nxb, nyb);
  for (int p = 0; p < fixc; ++p) {
    int cx = -1, cy = -1;
    nyq_fix_decode(p, nlx, nly, nxb, nyb, &cx, &cy);
    CHECK(cx >= 0 && cx < nxb && cy >= 0 && cy < nyb, "fix pos out of range p=%d (%d,%d)", p, cx, cy);
    zcount[(size_t)cx * nyb + cy]++;
    const int n1 = nyq_index_at(cx, cy, nlx, nly, nxb, nyb);
    const int n2 = nyq_index_at(compact_mirror(cx, nxb), compact_mirror(cy, nyb), nlx, nly, nxb, nyb);
    CHECK(n1 < nlx + nly && n2 < nlx + nly, "fix nyq index");
  }
  for (size_t k = 0; k < zcount.size(); ++k)
    CHECK(zcount[k] == 1, "Z position %zu written %d times (nlx=%d nly=%d nxe=%d nye=%d)", k,
          zcount[k], nlx, nly, nxe, nye);
  // compact -> padded injective and in range
  for (int ax = 0; ax < 2; ++ax) {
    const int nb = ax ? nyb : nxb, ne = ax ? nye : nxe;
    std::set<int> seen;
    for (int c = 0; c < nb; ++c) {
      const int p = padded_of_compact(c, nb, ne);
      CHECK(p >= 0 && p < ne, "padded out of range");
      CHECK(seen.insert(p).second, "padded collision");
    }
  }
  CHECK(idle <= nlx, "too many idle lanes");
}

int main() {
  int cases = 0;
  const int nls[] = {2, 4, 6, 8, 10, 16, 32, 64, 128};
  for (int nlx : nls)
    for (int nly : nls)
      for (int ex = 0; ex <= 6; ex += 2)
        for (int ey = 0; ey <= 6; ey += 2) {
          check_case(nlx, nly, nlx + ex, nly + ey);
          ++cases;
        }
  // odd clamped modes (nl == ne odd)
  const int odd[] = {1, 3, 5, 7, 9, 63};
  for (int a : odd)
    for (int b : odd) {
      check_case(a, b, a, b);
      ++cases;
    }
  for (int a : odd)
    for (int nly : nls) {
      check_case(a, nly, a, nly + 2);
      check_case(nly, a, nly + 4, a);
      cases += 2;
    }
  // four-step splits of long transforms: n1 * n2 = n, both within the limit,
  // n1 the smallest divisor at or above sqrt(n); -1 exactly when no divisor
  // pair fits
  const int limit = 10033;
  for (int n = 1; n <= 300000; ++n) {
    const int s = fft_split(n, limit);
    if (n <= limit) {
      CHECK(s == 0, "n %d fits one workgroup but split %d", n, s);
      continue;
    }
    bool any = false;
    int best = -1;
    for (int d = 2; d <= limit; ++d)
      if (n % d == 0 && n / d <= limit) {
        any = true;
        if ((int64_t)d * d >= n && best < 0) best = d;
      }
    if (!any) {
      CHECK(s == -1, "n %d has no split but got %d", n, s);
    } else {
      CHECK(s > 0 && n % s == 0 && s <= limit && n / s <= limit && s == best, "n %d split %d (best %d)", n, s, best);
    }
  }
  ++cases;
  // production shapes
  check_case(512, 512, 1536, 1536);
  check_case(64, 128, 960, 384);
  check_case(64, 32, 192, 544);
  check_case(128, 64, 128, 64);
  cases += 4;
  printf("OK %d\n", cases);
  return fails ? 1 : 0;
}
