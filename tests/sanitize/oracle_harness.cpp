// Test infrastructure: reads a .data/.grid pair with the library's host parsers
// (mphexplicit_amd/csrc/mph_io.cpp) and steps the CPU oracle (oracle/mph_oracle.c), so that
// both can be built with -fsanitize=address,undefined and run on the CPU.
//   oracle_harness <data> <grid> <dim> <steps>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/mph.h"
#include "../../oracle/mph_oracle.h"

int main(int argc, char** argv)
{
    if (argc < 5) return 2;
    mph_params p;
    mph_params_default(&p);
    if (mph_read_data_file(argv[1], &p)) return 3;
    int n = 0;
    if (mph_read_grid_header(argv[2], &p, &n)) return 4;
    std::vector<int> type(n);
    std::vector<double> x(3 * (size_t)n), v(3 * (size_t)n);
    if (mph_read_grid_particles(argv[2], n, type.data(), x.data(), v.data())) return 5;
    p.dim = std::atoi(argv[3]);
    orc_state* s = orc_create(&p, n, type.data(), x.data(), v.data(), 1);
    if (!s) return 6;
    const int rb = orc_step(s, std::atoi(argv[4]));
    std::vector<double> vp(n);
    orc_virial(s, vp.data());
    std::vector<long> ptr(orc_count(s) + 1);
    std::vector<int> ind(orc_neighbor_total(s) > 0 ? orc_neighbor_total(s) : 1);
    orc_neighbors(s, ptr.data(), ind.data());
    std::printf("n %d rebuilds %d time %.17g pairs %ld\n", n, rb, orc_time(s), ptr.back());
    orc_destroy(s);
    return 0;
}
