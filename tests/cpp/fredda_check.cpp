// fredda_check.cpp -- reads a FREDDA candidate file with the drop-in
// clusterization::countCandidatesInFreddaFile / readFreddaCandFile
// (Clusterization.cu:362-415) and writes it back as the CSV of
// writeCandHostArrayToCSV (Clusterization.cu:338-358).  Host code only.
//   usage: fredda_check <in.cand> <out.csv>    (prints the count)
#include <cstdio>
#include <vector>

#include "sps/search.hpp"

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const int n = clusterization::countCandidatesInFreddaFile(argv[1]);
    std::vector<Cand> c((size_t)n);
    clusterization::readFreddaCandFile(argv[1], c.data(), (size_t)n);
    clusterization::writeCandHostArrayToCSV(c.data(), n, argv[2]);
    printf("%d\n", n);
    return 0;
}
