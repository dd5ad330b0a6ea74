// Command-line front end over the C++ host classes: the flows of the
// reference's T-BLAEQ executable (main.cu:35-165) -- build an index from a
// dataset file, generate a synthetic index, or run a query file against a
// saved index and append the timing row to a CSV -- with the same flags,
// defaults and checks.  The argument parser is this file's own.
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/tblaeq_host.hpp"

using namespace tblaeq_host;

namespace {

struct Options {
    bool build = false, query = false, gen = false;
    std::string dataset, indexPath = "indexes/", queryFile, resultPath = "tempResult.csv";
    int maxQueries = 10, queryType = 0, k = 10;
    bool forceCpu = false;
    size_t height = 4;
    std::string ratios = "100,50,20";
    size_t genN = 1000000, genD = 3;
    double genMin = 1.0, genMax = 100.0;
    bool genInt = false;
    uint64_t genSeed = 12345;
    double genSigma = 3.0;
    std::string forcePolicy;
};

const char* kUsage =
    "usage: tblaeq-cli (--build-index | --test-query | --gen-index) [options]\n"
    "  --build-index            build an index from --dataset and save it to --index-path\n"
    "  --test-query             run --query-file against the index at --index-path\n"
    "  --gen-index              generate a synthetic hierarchy and save it to --index-path\n"
    "  -d, --dataset PATH       dataset text file ('N D' then N rows of D numbers)\n"
    "  -i, --index-path DIR     index directory (default indexes/)\n"
    "  -f, --query-file PATH    query text file\n"
    "  -r, --result-path PATH   CSV the query timings are appended to (default tempResult.csv)\n"
    "  -q, --max-queries N      at most N queries of the file (default 10)\n"
    "  -t, --query-type T       0 = range, 1 = KNN (default 0)\n"
    "  -k, --knn-k K            K of KNN queries (default 10)\n"
    "  --force-cpu BOOL         build the index with the CPU k-means (default false)\n"
    "  --height H               levels of the hierarchy, at least 2 (default 4)\n"
    "  --ratios R1,R2,...       coarsening ratios, height - 1 of them (default 100,50,20)\n"
    "  --gen-N N --gen-D D --gen-min A --gen-max B --gen-int --gen-seed S --gen-sigma-divisor V\n"
    "                           synthetic hierarchy (defaults 1000000, 3, 1.0, 100.0, off, 12345, 3.0)\n"
    "  --force-policy L0|L1|L2|L3  one memory policy for every level (default: chosen from free memory)\n";

[[noreturn]] void usageError(const std::string& msg) {
    std::cerr << "tblaeq-cli: " << msg << "\n" << kUsage;
    std::exit(2);
}

bool parseBool(const std::string& v) {
    if (v == "1" || v == "true" || v == "True" || v == "TRUE" || v == "on" || v == "yes") return true;
    if (v == "0" || v == "false" || v == "False" || v == "FALSE" || v == "off" || v == "no") return false;
    usageError("not a boolean: '" + v + "'");
}

template <class T>
T parseNumber(const std::string& flag, const std::string& v) {
    std::istringstream is(v);
    T x{};
    if (!(is >> x) || !is.eof()) usageError("bad value for " + flag + ": '" + v + "'");
    return x;
}

Options parse(int argc, char** argv) {
    Options o;
    std::map<std::string, std::string> alias = {{"-d", "--dataset"},       {"-i", "--index-path"},
                                                {"-f", "--query-file"},    {"-r", "--result-path"},
                                                {"-q", "--max-queries"},   {"-t", "--query-type"},
                                                {"-k", "--knn-k"}};
    int modes = 0;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i], v;
        if (a == "-h" || a == "--help") {
            std::cout << kUsage;
            std::exit(0);
        }
        const size_t eq = a.find('=');
        bool inlineValue = false;
        if (a.rfind("--", 0) == 0 && eq != std::string::npos) {
            v = a.substr(eq + 1);
            a = a.substr(0, eq);
            inlineValue = true;
        }
        if (alias.count(a)) a = alias[a];
        auto value = [&]() -> std::string {
            if (inlineValue) return v;
            if (i + 1 >= argc) usageError(a + " needs a value");
            return argv[++i];
        };
        if (a == "--build-index") o.build = true, ++modes;
        else if (a == "--test-query") o.query = true, ++modes;
        else if (a == "--gen-index") o.gen = true, ++modes;
        else if (a == "--gen-int") o.genInt = inlineValue ? parseBool(v) : true;
        else if (a == "--dataset") o.dataset = value();
        else if (a == "--index-path") o.indexPath = value();
        else if (a == "--query-file") o.queryFile = value();
        else if (a == "--result-path") o.resultPath = value();
        else if (a == "--max-queries") o.maxQueries = parseNumber<int>(a, value());
        else if (a == "--query-type") o.queryType = parseNumber<int>(a, value());
        else if (a == "--knn-k") o.k = parseNumber<int>(a, value());
        else if (a == "--force-cpu") o.forceCpu = parseBool(value());
        else if (a == "--height") o.height = parseNumber<size_t>(a, value());
        else if (a == "--ratios") o.ratios = value();
        else if (a == "--gen-N") o.genN = parseNumber<size_t>(a, value());
        else if (a == "--gen-D") o.genD = parseNumber<size_t>(a, value());
        else if (a == "--gen-min") o.genMin = parseNumber<double>(a, value());
        else if (a == "--gen-max") o.genMax = parseNumber<double>(a, value());
        else if (a == "--gen-seed") o.genSeed = parseNumber<uint64_t>(a, value());
        else if (a == "--gen-sigma-divisor") o.genSigma = parseNumber<double>(a, value());
        else if (a == "--force-policy") o.forcePolicy = value();
        else usageError("unknown argument '" + a + "'");
    }
    if (modes != 1) usageError("exactly one of --build-index, --test-query, --gen-index is required");
    return o;
}

// Comma-separated positive ratios (main.cu:11-29 checks the same conditions).
std::vector<size_t> parseRatios(const std::string& s) {
    std::vector<size_t> out;
    size_t b = 0;
    while (b <= s.size()) {
        const size_t e = s.find(',', b);
        const std::string tok = s.substr(b, e == std::string::npos ? std::string::npos : e - b);
        if (tok.empty()) throw std::invalid_argument("--ratios '" + s + "' has an empty entry");
        size_t used = 0;
        const unsigned long long r = std::stoull(tok, &used);
        if (used != tok.size() || r == 0) throw std::invalid_argument("--ratios entries must be positive integers");
        out.push_back(static_cast<size_t>(r));
        if (e == std::string::npos) break;
        b = e + 1;
    }
    return out;
}

int run(const Options& o) {
    if (o.height < 2) throw std::runtime_error("--height must be at least 2");
    const std::vector<size_t> ratios = parseRatios(o.ratios);
    if (ratios.size() != o.height - 1) throw std::runtime_error("--ratios needs exactly --height - 1 entries");
    if (o.build) {
        QueryHandler h(o.forceCpu, o.dataset, o.height, ratios);
        h.saveIndex(o.indexPath);
        std::cout << "index of " << h.getSize() << " points saved to " << o.indexPath << "\n";
    }
    if (o.gen) {
        if (!(o.genSigma > 0.0)) throw std::runtime_error("--gen-sigma-divisor must be positive");
        QueryHandler h(o.genN, o.genD, o.genMin, o.genMax, o.genInt, o.height, ratios, o.genSeed, o.genSigma);
        h.saveIndex(o.indexPath);
        std::cout << "synthetic index of " << h.getSize() << " points saved to " << o.indexPath << "\n";
    }
    if (o.query) {
        if (o.queryType != 0 && o.queryType != 1)
            throw std::runtime_error("--query-type must be 0 (range) or 1 (KNN), not " + std::to_string(o.queryType));
        QueryHandler h(o.indexPath, true);
        const QueryType t = o.queryType == 1 ? QueryType::POINT : QueryType::RANGE;
        if (!o.forcePolicy.empty()) {
            const LevelPolicy lp = IndexPolicy::parseLevel(o.forcePolicy);
            std::cout << "every level runs under policy " << levelPolicyName(lp) << "\n";
            h.prepareForQuery(lp);
        } else {
            h.prepareForQuery();
        }
        QueryResult r = h.performQuery(o.queryFile, t, false, o.maxQueries, static_cast<size_t>(o.k));
        saveQueryResult(r, o.resultPath);
        std::cout << "timings appended to " << o.resultPath << "\n";
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    const Options o = parse(argc, argv);
    try {
        return run(o);
    } catch (const std::exception& e) {
        std::cerr << "tblaeq-cli: " << e.what() << "\n";
        return 1;
    }
}
