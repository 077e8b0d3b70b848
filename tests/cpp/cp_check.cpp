// Prints Clopper-Pearson bounds of drone-race_amd/csrc/dr_cp.h (host build)
// for the (k, n, alpha) triples read from stdin, one per line:
// "k n alpha" -> "lower upper".
#include <cstdio>

#include "../../drone-race_amd/csrc/dr_cp.h"

int main()
{
    unsigned k, n;
    double alpha;
    while (std::scanf("%u %u %lf", &k, &n, &alpha) == 3)
        std::printf("%.17g %.17g\n", dr::cp::cp_lower(k, n, alpha), dr::cp::cp_upper(k, n, alpha));
    return 0;
}
