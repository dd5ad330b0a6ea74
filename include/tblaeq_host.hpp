// Host-side C++ mirror of the reference's query interface, MI355X-native inside.
//
// Same class names, argument meaning and error behaviour as the reference
// (src/core/*.cuh, src/Data_Structures/*.cuh), with two deliberate changes:
//  - device errors are reported as exceptions / status codes instead of
//    std::exit (check.cuh:15-25);
//  - level kernels keep element counts on the device, so a query issues its
//    whole level loop without host synchronisation and the engine reads every
//    level's size once at the end of the query.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <chrono>
#include <climits>
#include <iosfwd>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "tblaeq.h"
#include "tblaeq_build.h"

namespace tblaeq_host {

// A failed HIP call or library entry point; `status` is the TBLAEQ_ERR_* code
// a C caller receives (TBLAEQ_ERR_NOMEM for a failed device allocation, the
// entry point's own code when a C-ABI call failed).
struct HipError : std::runtime_error {
    explicit HipError(const std::string& m, int s = TBLAEQ_ERR_HIP) : std::runtime_error(m), status(s) {}
    int status;
};

// An index, dataset or query file that is missing, unreadable or malformed
// (the reference throws std::runtime_error for these, File.cu and
// IndexSerializer.cu); C callers receive TBLAEQ_ERR_IO.
struct IoError : std::runtime_error {
    explicit IoError(const std::string& m) : std::runtime_error(m) {}
};

// A call that does not fit the object's state, e.g. reading a point from a
// batch of range queries (the reference throws std::runtime_error, File.cu);
// C callers receive TBLAEQ_ERR_INVALID.
struct UsageError : std::runtime_error {
    explicit UsageError(const std::string& m) : std::runtime_error(m) {}
};

void hipCheck(hipError_t e, const char* what);
void tbCheck(int rc, const char* what);

// Status code of the exception being handled (call only inside a catch
// block), with its message recorded as the thread's last error:
// HipError -> its status, IoError -> TBLAEQ_ERR_IO, std::bad_alloc ->
// TBLAEQ_ERR_NOMEM, UsageError and std::logic_error (invalid_argument,
// out_of_range, a strategy used before it was prepared) ->
// TBLAEQ_ERR_INVALID, anything else -> TBLAEQ_ERR_INTERNAL.
int statusOfCurrentException(const char* what) noexcept;

// RAII device buffer (Memory.cuh:33-168).
template <class T>
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t n) { resize(n); }
    DeviceBuffer(const T* host, size_t n) { reset(host, n); }
    ~DeviceBuffer() { free(); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_), n_(o.n_) {
        o.p_ = nullptr;
        o.n_ = 0;
    }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        if (this != &o) {
            free();
            p_ = o.p_;
            n_ = o.n_;
            o.p_ = nullptr;
            o.n_ = 0;
        }
        return *this;
    }
    void resize(size_t n) {
        if (n == n_ && p_) return;
        free();
        if (n) hipCheck(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)), "hipMalloc");
        n_ = n;
    }
    void reset(const T* host, size_t n) {
        resize(n);
        if (n) hipCheck(hipMemcpy(p_, host, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
    }
    void free() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    // Copies n elements to the host (the caller has synchronised the stream
    // that wrote them).
    void download(T* host, size_t n) const {
        if (n) hipCheck(hipMemcpy(host, p_, n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    }
    T* data() const { return p_; }
    size_t size() const { return n_; }
    size_t bytes() const { return n_ * sizeof(T); }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

// ---- data structures (Data_Structures.cuh, File.cuh) ----------------------

enum class QueryType { RANGE = 0, POINT = 1 };

struct PointCloud {  // Data_Structures.cuh:35-114 (AoS doubles)
    int64_t size = 0;
    int64_t dim = 0;
    std::vector<double> data;
};

// One level's point set: ids[nnz] and AoS vals[nnz*dim], host or device,
// owning or not (Data_Structures.cuh:134-229).  dCount, when set, is the device
// word holding nnz; nnz_ is then valid after QueryEngine synchronises.
class SparseGrid {
public:
    SparseGrid() = default;
    SparseGrid(size_t numRows, size_t dim, size_t nnz, uint64_t* ids, double* vals)
        : numRows_(numRows), dim_(dim), nnz_(nnz), ids_(ids), vals_(vals) {}
    size_t get_num_rows() const { return numRows_; }
    size_t get_dimensions() const { return dim_; }
    size_t get_nnz_nums() const { return nnz_; }
    void set_nnz(size_t n) { nnz_ = n; }
    const uint64_t* get_ids_() const { return ids_; }
    const double* get_vals_() const { return vals_; }
    uint64_t* ids() { return ids_; }
    double* vals() { return vals_; }
    void set_ids(uint64_t* p) { ids_ = p; }
    void set_vals(double* p) { vals_ = p; }
    const uint64_t* dCount = nullptr;  // device-resident nnz (MI355X path)
    uint64_t nnzBound = 0;             // host upper bound on nnz while dCount is pending

private:
    size_t numRows_ = 0, dim_ = 0, nnz_ = 0;
    uint64_t* ids_ = nullptr;
    double* vals_ = nullptr;
};
using GridAsSparseMatrix = SparseGrid;

// CSC prolongation tensor (Data_Structures.cuh:299-390).
struct SparseTensorCscFormat {
    uint64_t D = 0, rows = 0, cols = 0, nnz = 0;
    bool isAos = true;
    std::vector<uint64_t> rowIds, nnzPerCol, colRes;
    std::vector<double> vals;
    size_t get_row_nums() const { return rows; }
    size_t get_col_nums() const { return cols; }
    size_t get_nnz_nums() const { return nnz; }
    size_t get_dim() const { return D; }
};

// Query batch (File.cuh:52-109): range rows hold lo then hi (2*dim values).
struct Query {
    Query() = default;
    Query(int len, int d, QueryType t) : length(len), dim(d), type(t), data(static_cast<size_t>(len) * d * (t == QueryType::RANGE ? 2 : 1)) {}
    int length = 0;
    int dim = 0;
    QueryType type = QueryType::RANGE;
    std::vector<double> data;
    std::string queryRangeInfo;
    std::vector<double> getQueryPoint(int i) const;
    std::pair<std::vector<double>, std::vector<double>> getQueryRange(int i) const;
    void setQueryPoint(int i, const std::vector<double>& p);
    void setQueryRange(int i, const std::vector<double>& lo, const std::vector<double>& hi);
};

PointCloud loadFromFile(const std::string& path);                 // File.cu:8-50
Query loadQueryPointFromFile(const std::string& path);            // File.cu:210
Query loadQueryRangeFromFile(const std::string& path);            // File.cu:212-216
void saveQueryToFile(const Query& q, const std::string& path);    // File.cu:218-247

// ---- memory policy (MemoryPolicy.cuh) --------------------------------------

enum class LevelPolicy { L0 = 0, L1 = 1, L2 = 2, L3 = 3 };
const char* levelPolicyName(LevelPolicy p);

struct IndexData;

struct IndexPolicy {
    std::vector<LevelPolicy> levels;
    LevelPolicy operator[](size_t l) const { return levels[l]; }
    bool anyL0() const;
    bool anyStreaming() const;
    static IndexPolicy uniform(size_t n, LevelPolicy p) {
        IndexPolicy r;
        r.levels.assign(n, p);
        return r;
    }
    static LevelPolicy parseLevel(const std::string& s);
};

struct DeviceMemoryInfo {
    size_t freeMem = 0, totalMem = 0;
    static DeviceMemoryInfo query();
};

struct MemoryStats {  // IndexData.cuh:43-70
    size_t hostPTensors = 0, hostMaps = 0, hostRadius = 0, hostCoarsestMesh = 0, hostTotal = 0;
    size_t devicePTensorVals = 0, deviceMaps = 0, deviceRadius = 0, deviceCoarsestMesh = 0, deviceSpTSpMBufs = 0,
           deviceTotal = 0, grandTotal = 0;
    void print(std::ostream& os) const;
};

class IQueryStrategy;

class PolicyScheduler {
public:
    static constexpr double kSafetyFactor = 0.85;
    // Greedy per-level policy, coarsest first (MemoryPolicy.cu:100-177).  With
    // countWorkBuffers the L0 cost also counts the level's pre-allocated work
    // buffers (the reference leaves them out, IndexData.cu:114-136).
    static IndexPolicy recommend(const IndexData& idx, size_t freeMem, bool countWorkBuffers = true);
    static IndexPolicy recommend(const IndexData& idx);
    static void printReport(const IndexData& idx, const IndexPolicy& p, std::ostream& os);
    static std::unique_ptr<IQueryStrategy> make(IndexData& idx, const IndexPolicy& policy);
    static std::unique_ptr<IQueryStrategy> makeForLevel(IndexData& idx, size_t l);
    static size_t l0WorkBytes(const IndexData& idx, size_t l);
};

// ---- index (IndexData.cuh) -------------------------------------------------

struct LevelDevice {
    DeviceBuffer<uint64_t> colRes, nnzPerCol, childIds;
    DeviceBuffer<double> pvals, radius;
    uint64_t emptyCols = 0;
    bool resident = false;  // pvals/radius resident (L0/L1)
};

// Per-level L0 work buffers: survivors and prolongation output (pre-allocated,
// reused by every query; the L0 path performs no allocation).
struct LevelWork {
    DeviceBuffer<uint64_t> survIds, survOff, outIds;
    DeviceBuffer<double> survVals, outVals;
    // learned KNN radius: survivors within the proposed radius (last level)
    DeviceBuffer<uint64_t> radIds, radOff;
    DeviceBuffer<double> radVals;
};

struct IndexData {
    size_t D = 0, N = 0, height = 0, intervals = 0;
    bool isAosArch = true;
    std::string datasetName;
    std::vector<std::unique_ptr<SparseTensorCscFormat>> pTensors;  // [intervals], 0 = coarsest
    std::vector<std::vector<double>> meshMaxRadius;                 // [intervals]
    std::vector<size_t> meshSizes;                                  // finest first
    std::vector<size_t> ratios;
    std::vector<std::vector<uint64_t>> maps;  // [height], maps[0] empty
    size_t coarsestRows = 0;
    std::vector<uint64_t> coarsestIds;
    std::vector<double> coarsestVals;
    std::vector<double> finestData;  // generator output (finest layer), when generated
    std::vector<uint64_t> buildStats;  // builds: per level (finest first) the six tblaeq_kmeans_ex stats words

    // device side
    tblaeq_ctx* ctx = nullptr;
    std::vector<LevelDevice> dLevels;
    std::vector<LevelWork> work;
    DeviceBuffer<uint64_t> dCoarsestIds;
    DeviceBuffer<double> dCoarsestVals;
    DeviceBuffer<uint64_t> dCounts;  // per level: [survivors, outputs] + coarsest count
    DeviceBuffer<double> dQuery;     // lo|hi or query point (uploaded once per query)
    double* hQuery = nullptr;        // pinned staging for dQuery (2D values + the learned radius)
    double learnedRadius = 0.0;      // > 0: the last KNN level keeps only survivors with lb <= this
    IndexPolicy activePolicy;
    bool permanentDataOnDevice = false;
    bool workReady = false;
    std::vector<size_t> hostSurv;  // survivors of the last query per level (host-synchronising strategies)
    MemoryStats stats;

    IndexData() = default;
    ~IndexData();
    IndexData(const IndexData&) = delete;
    IndexData& operator=(const IndexData&) = delete;

    void computeStats();
    void ensureCtx();
    void uploadPermanentData();                    // maps (as child ids), CSC structure, coarsest mesh
    void uploadLevelValues(size_t l);              // P values + radius resident
    void allocLevelWork(size_t l);                 // L0 work buffers
    tblaeq_level levelDesc(size_t l) const;
    size_t levelCols(size_t l) const { return pTensors[l]->cols; }
    size_t levelNnz(size_t l) const { return pTensors[l]->nnz; }
};

class IndexSerializer {  // IndexSerializer.cu:11-259, byte-compatible
public:
    static void save(const IndexData& d, const std::string& dir);
    static IndexData* load(const std::string& dir);
};

class IndexBuilder {  // IndexBuilder.cu:14-234
public:
    static const std::vector<size_t> kDefaultRatios;
    static constexpr size_t kDefaultHeight = 4;
    static size_t computeCentroidCount(size_t n, size_t ratio);
    // GPU build; data32 (when data is null) holds FP32 values, converted to
    // FP64 on the device; opt null = the CPU build's settings.
    static IndexData* buildGpu(const double* data, size_t N, size_t D, const std::string& name, size_t height,
                               const std::vector<size_t>& ratios, const tblaeq_kmeans_options* opt = nullptr,
                               const float* data32 = nullptr);
    static IndexData* build(const double* data, size_t N, size_t D, const std::string& name, bool forceUseCPU,
                            size_t height, const std::vector<size_t>& ratios, const tblaeq_kmeans_options* opt = nullptr);
    static IndexData* buildRandom(size_t N, size_t D, double valMin, double valMax, bool isInt,
                                  const std::string& name, size_t height, const std::vector<size_t>& ratios,
                                  uint64_t seed, double sigmaDivisor);
    // Assembles P-tensors, maps, radii and the coarsest mesh from layers
    // (finest first) and labels[i] (layer i -> layer i+1).
    static IndexData* assemble(size_t N, size_t D, const std::vector<size_t>& ratios,
                               std::vector<std::vector<double>>& layers,
                               const std::vector<std::vector<uint64_t>>& labels);
    // build / buildGpu after the reference's plan adjustment (a level that
    // would collapse to one centroid ends the hierarchy, QueryHandler.cu:23-45),
    // with FP64 data or, when data is null, FP32 data32; no reference
    // counterpart for the options.
    static IndexData* buildPlanned(const double* data, const float* data32, size_t N, size_t D,
                                   const std::string& name, bool forceUseCPU, size_t height,
                                   const std::vector<size_t>& ratios, const tblaeq_kmeans_options* opt);
    // No reference counterpart: the generator's hierarchy shape built on
    // `threads` host threads (0 = all) from per-chunk streams, for indexes too
    // large for the single-stream generator; not the reference's sequence.
    static IndexData* buildSynthetic(size_t N, size_t D, double valMin, double valMax, bool isInt,
                                     const std::string& name, size_t height, const std::vector<size_t>& ratios,
                                     uint64_t seed, double sigmaDivisor, unsigned threads);
    // assemble() on host threads, each column's rows in ascending point order.
    static IndexData* assembleStable(size_t N, size_t D, const std::vector<size_t>& ratios,
                                     std::vector<std::vector<double>>& layers,
                                     const std::vector<std::vector<uint64_t>>& labels, unsigned threads);
};

// ---- strategies (IQueryStrategy.cuh) ----------------------------------------

struct LevelResult {
    SparseGrid* grid = nullptr;
    bool ownsIds = true;
    bool ownsVals = true;
    bool onHost = false;
};

class IQueryStrategy {
public:
    virtual ~IQueryStrategy() = default;
    virtual void prepare(std::ostream& reportOs) = 0;
    virtual LevelResult runLevel(size_t l, SparseGrid* currentGrid, QueryType qType, const double* lo,
                                 const double* hi, const double* queryPoint, size_t K) = 0;
};

class L0Strategy : public IQueryStrategy {
public:
    explicit L0Strategy(IndexData& idx) : idx_(idx) {}
    void prepare(std::ostream& os) override;
    LevelResult runLevel(size_t l, SparseGrid* cur, QueryType t, const double* lo, const double* hi,
                         const double* q, size_t K) override;

private:
    IndexData& idx_;
};

class L1Strategy : public IQueryStrategy {
public:
    explicit L1Strategy(IndexData& idx) : idx_(idx) {}
    void prepare(std::ostream& os) override;
    LevelResult runLevel(size_t l, SparseGrid* cur, QueryType t, const double* lo, const double* hi,
                         const double* q, size_t K) override;

private:
    IndexData& idx_;
};

class L2Strategy : public IQueryStrategy {
public:
    explicit L2Strategy(IndexData& idx) : idx_(idx) {}
    void prepare(std::ostream& os) override;
    LevelResult runLevel(size_t l, SparseGrid* cur, QueryType t, const double* lo, const double* hi,
                         const double* q, size_t K) override;

private:
    IndexData& idx_;
};

class L3Strategy : public IQueryStrategy {
public:
    explicit L3Strategy(IndexData& idx, size_t tileBudget = 0) : idx_(idx), tileBudget_(tileBudget) {}
    void prepare(std::ostream& os) override;
    LevelResult runLevel(size_t l, SparseGrid* cur, QueryType t, const double* lo, const double* hi,
                         const double* q, size_t K) override;
    static size_t forcedTileBytes;  // testing hook: fixed device budget for tiles (0 = from free memory)

private:
    IndexData& idx_;
    size_t tileBudget_;
};

void freeLevelResult(const LevelResult& r);

// ---- engine (QueryEngine.cuh) -------------------------------------------------

struct QueryResult {
    ~QueryResult();
    QueryResult() = default;
    QueryResult(const QueryResult&) = delete;
    QueryResult& operator=(const QueryResult&) = delete;
    QueryResult(QueryResult&&) = default;

    QueryType type = QueryType::RANGE;
    int errorCode = 0;
    long totalTimeUs = 0;
    size_t queryCount = 0;
    std::string datasetName, queryParam, memoryPolicy;
    size_t datasetSize = 0, datasetDim = 0;
    std::vector<double> queryRangeVolume;
    std::vector<size_t> fineMeshSize;
    std::vector<size_t> levelOriginalSize;
    std::vector<std::vector<size_t>> levelFineMeshSize;
    std::vector<std::vector<size_t>> levelSurvivors;  // clusters kept by each level's prune
    std::vector<long> queryTimeUs;
    std::vector<double> queryTimeUsExact;  // sub-microsecond timings (same clock)
    std::vector<SparseGrid*> fineMesh;     // host copies when saveFineMesh
    // exact refine (new): per query sorted hit ids (range) or top-K ids + squared distances (KNN)
    std::vector<size_t> refinedCount;
    std::vector<std::vector<uint64_t>> refinedIds;
    std::vector<std::vector<double>> refinedDist2;
};

class QueryEngine {
public:
    struct RunConfig {
        bool saveFineMesh = false;
        int maxQueryCount = std::numeric_limits<int>::max();
        size_t K = 0;
        bool refine = false;   // run the exact refine after the traversal
        bool copyRefined = true;  // copy refined ids/distances to the host (else counts only)
        bool quiet = true;     // no console output inside the query loop
        // learned KNN radius per query (refine must be on): the last level
        // prolongs only survivors within the radius; a query whose K-th result
        // lies beyond it (or that finds fewer than K) retries with a larger
        // radius, then runs in full
        const double* radii = nullptr;
        uint64_t* fallbacks = nullptr;
    };
    static QueryResult run(IndexData& idx, IQueryStrategy& strategy, const Query& q, const RunConfig& cfg);
    // One query through the level loop; fills levelSizes and returns the final grid.
    static LevelResult runSingleQuery(IndexData& idx, QueryType t, const double* lo, const double* hi,
                                      const double* qp, size_t K, size_t* levelFineMeshSize,
                                      size_t* levelSurvivors = nullptr);
};

void saveQueryResult(const QueryResult& r, const std::string& path);  // QueryEngine.cu:254-305

// ---- facade (QueryHandler.cuh) ---------------------------------------------------

class QueryHandler {
public:
    explicit QueryHandler(bool forceUseCPU, const std::string& datasetPath,
                          size_t height = IndexBuilder::kDefaultHeight,
                          const std::vector<size_t>& ratios = IndexBuilder::kDefaultRatios);
    explicit QueryHandler(bool forceUseCPU, const PointCloud& dataset, const std::string& name = "in-memory",
                          size_t height = IndexBuilder::kDefaultHeight,
                          const std::vector<size_t>& ratios = IndexBuilder::kDefaultRatios);
    QueryHandler(size_t N, size_t D, double valMin, double valMax, bool isInt,
                 size_t height = IndexBuilder::kDefaultHeight,
                 const std::vector<size_t>& ratios = IndexBuilder::kDefaultRatios, uint64_t seed = 12345,
                 double sigmaDivisor = 3.0, const std::string& name = "random");
    QueryHandler(const std::string& indexPath, bool loadFromIndex);
    // Takes over an index built elsewhere (e.g. IndexBuilder::buildSynthetic);
    // no reference counterpart.
    explicit QueryHandler(std::unique_ptr<IndexData> index);

    void saveIndex(const std::string& dir) const;
    void prepareForQuery();
    void prepareForQuery(IndexPolicy policy);
    void prepareForQuery(LevelPolicy level);
    void prepareForQuery(size_t freeMemOverride);  // scheduler with an injected free-memory value
    const MemoryStats& getMemoryStats() const { return idx_->stats; }
    const IndexPolicy& getActivePolicy() const { return idx_->activePolicy; }
    QueryResult performQuery(const std::string& queryPath, QueryType qType, bool saveFineMesh = false,
                             int maxQueryCount = std::numeric_limits<int>::max(), size_t K = 0);
    QueryResult performQuery(const Query& q, const QueryEngine::RunConfig& cfg);
    QueryResult performSingleKNNQuery(const std::vector<double>& queryPoint, size_t k, bool saveFineMesh = false);
    QueryResult performSingleRangeQuery(const std::vector<double>& queryUpperBound,
                                        const std::vector<double>& queryLowerBound, bool saveFineMesh = false);
    size_t getSize() const { return idx_->N; }
    size_t getDim() const { return idx_->D; }
    IndexData& index() { return *idx_; }
    bool prepared() const { return prepared_; }

private:
    std::unique_ptr<IndexData> idx_;
    std::unique_ptr<IQueryStrategy> strategy_;
    bool prepared_ = false;
    void doPrepare(const IndexPolicy& policy);
    void ensurePrepared();
};

}  // namespace tblaeq_host
