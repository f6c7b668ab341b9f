// Test infrastructure (oracle side): prints the reference's runtime tables.
// Links against the reference's own table sources, compiled unmodified from
// /root/reference by oracle/Makefile into oracle/_ref/.  Output: one JSON
// object {name: nested list of hex doubles}, orders 0..9 zero-padded to 10,
// in the same shape as exahype_amd/tables/dg_tables.json.
#include <cstdio>
#include <set>
#include "kernels/GaussLegendreQuadrature.h"
#include "kernels/GaussLobattoQuadrature.h"
#include "kernels/DGMatrices.h"

static void hexd(double v) { std::printf("\"%a\"", v); }

static void vec2(const char* name, double** t, bool first) {
  std::printf("%s\"%s\":[", first ? "" : ",", name);
  for (int n = 0; n < 10; ++n) {
    std::printf(n ? ",[" : "[");
    for (int i = 0; i < 10; ++i) { if (i) std::printf(","); hexd(i <= n ? t[n][i] : 0.0); }
    std::printf("]");
  }
  std::printf("]");
}

static void mat3(const char* name, double*** t) {
  std::printf(",\"%s\":[", name);
  for (int n = 0; n < 10; ++n) {
    std::printf(n ? ",[" : "[");
    for (int i = 0; i < 10; ++i) {
      std::printf(i ? ",[" : "[");
      for (int j = 0; j < 10; ++j) { if (j) std::printf(","); hexd(i <= n && j <= n ? t[n][i][j] : 0.0); }
      std::printf("]");
    }
    std::printf("]");
  }
  std::printf("]");
}

int main() {
  std::set<int> orders;
  for (int n = 0; n < 10; ++n) orders.insert(n);
  kernels::initGaussLegendreNodesAndWeights(orders);
  kernels::initGaussLobattoNodesAndWeights(orders);
  kernels::initDGMatrices(orders);
  std::printf("{");
  vec2("gl_weights", kernels::gaussLegendreWeights, true);
  vec2("gl_nodes", kernels::gaussLegendreNodes, false);
  vec2("gll_weights", kernels::gaussLobattoWeights, false);
  vec2("gll_nodes", kernels::gaussLobattoNodes, false);
  vec2("F0", kernels::F0, false);
  vec2("FLCoeff", kernels::FLCoeff, false);
  vec2("FRCoeff", kernels::FRCoeff, false);
  mat3("Kxi", kernels::Kxi);
  mat3("iK1", kernels::iK1);
  mat3("dudx", kernels::dudx);
  mat3("equidistantGridProjector1d", kernels::equidistantGridProjector1d);
  std::printf(",\"fineGridProjector1d\":[");
  for (int n = 0; n < 10; ++n) {
    std::printf(n ? ",[" : "[");
    for (int s = 0; s < 3; ++s) {
      std::printf(s ? ",[" : "[");
      for (int i = 0; i < 10; ++i) {
        std::printf(i ? ",[" : "[");
        for (int j = 0; j < 10; ++j) {
          if (j) std::printf(",");
          hexd(i <= n && j <= n ? kernels::fineGridProjector1d[n][s][i][j] : 0.0);
        }
        std::printf("]");
      }
      std::printf("]");
    }
    std::printf("]");
  }
  std::printf("]}\n");
  return 0;
}
