// Probe over the reference's femIntegrationRule (src/femIntegrationRule.cpp),
// compiled from the unmodified reference source into oracle/_ref/.  Test
// infrastructure only: it lets tests compare the oracle's Gauss tables with
// the reference's own getTotGP/getCoords/getWeights.
#include "femIntegrationRule.h"

extern "C" int ref_rule(int rule_type, int tot_nodes, int dims, double* coords /*[ngp][4]*/,
                        double* weights /*[ngp]*/) {
  try {
    femIntegrationRule r(static_cast<intRuleType>(rule_type));
    int n = r.getTotGP(tot_nodes, static_cast<elDim>(dims));
    femDoubleMat c = r.getCoords(tot_nodes, static_cast<elDim>(dims));
    femDoubleVec w = r.getWeights(tot_nodes, static_cast<elDim>(dims));
    for (int g = 0; g < n; g++) {
      for (int k = 0; k < 4; k++) coords[g * 4 + k] = (k < (int)c[g].size()) ? c[g][k] : 0.0;
      weights[g] = w[g];
    }
    return n;
  } catch (...) {
    return -1;
  }
}
