// A reference-side C++ call site (src/main.cu:128-161 of the reference) built
// against this library's host classes: generate an index, run a range query
// file through the per-query traversal and append the timing row to a CSV.
// Used by the tests: compiled and linked on the CPU, run on the GPU.
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "tblaeq_host.hpp"
using namespace tblaeq_host;

int main(int argc, char** argv) {
    if (argc != 3) {
        std::cerr << "usage: reference_caller QUERY_FILE RESULT_CSV\n";
        return 2;
    }
    const std::string queryPath = argv[1], resultPath = argv[2];
    const size_t N = 20000, D = 4, height = 3;
    const std::vector<size_t> ratios = {20, 10};
    QueryHandler h(N, D, 1.0, 100.0, false, height, ratios, 7, 3.0);  // QueryHandler.cu:101
    h.prepareForQuery();                                               // or (LevelPolicy::L0) / (IndexPolicy)
    QueryResult r = h.performQuery(queryPath, QueryType::RANGE, /*saveFineMesh=*/false);
    saveQueryResult(r, resultPath);                                    // QueryEngine.cu:254-305 columns
    std::cout << "queries " << r.queryCount << " fine mesh of the first " << r.fineMeshSize.at(0) << "\n";
    return 0;
}
