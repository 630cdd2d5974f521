/*
 * velodb_amd/operators.hpp -- C++ host side of the oblivious operator path,
 * header-only, over the C ABI of vo_api.h.
 *
 * It mirrors the reference's operator surface for this path so a caller of
 * VelODB's operators finds the same names, argument meaning and error
 * behaviour:
 *   TaskManager             include/cuda/task_manager.hpp:12-30
 *   ExecutionContext        include/catalog/execution_context.hpp:24 (getTaskManager)
 *   Result<T>               include/common/result.hpp (success / failure / value / error)
 *   AbstractOperator::next  include/operator/abstract_operator.hpp:18-36
 *   SortOperator            include/operator/sort_operator.hpp:15-19, src/operator/sort_operator.cpp:21-134
 *   FilterCompactionOperator include/operator/filter_compaction_operator.hpp:14-17,
 *                           src/operator/filter_compaction_operator.cpp:65-184
 *   SortMergeJoinOperator   include/operator/sort_merge_join_operator.hpp:14-33,
 *                           src/operator/sort_merge_join_operator.cpp:27-257
 *   JoinType                include/operator/join_type.hpp
 *   MaterializationOperator include/operator/materialization_operator.hpp,
 *                           src/operator/materialization_operator.cpp:362-517
 *   ObliviousTable          include/data/oblivious_table.hpp, src/data/oblivious_table.cpp:16-255
 *   ObliviousTableManager   include/data/oblivious_table_manager.hpp,
 *                           src/data/oblivious_table_manager.cpp:14-221
 *   AbstractPlanNode, SortPlanNode, FilterCompactionPlanNode, SortMergeJoinPlanNode,
 *   MaterializationPlanNode  include/planner/{sort,filter_compaction,sort_merge_join,
 *                           materialization}_plan_node.hpp and their .cpp files, and the
 *                           planner's wiring of them (src/planner/query_planner.cpp:247-252,
 *                           443-476, 518-544)
 * and adds PKFKJoinGroupSumOperator (and its plan node) for the oblivious
 * join + GROUP BY.
 *
 * Columns live in device memory with a power-of-two capacity (at least 4096
 * rows), as the reference's device ValueVector does (value_vector.hpp:43,
 * 248-272), so the sort and the compaction may use the pad rows in place.
 * A column may carry a null bitmap of one byte per row (non-zero = valid),
 * moved with its rows.
 *
 * Differences from the reference, by design of the oblivious path:
 *   - FilterCompactionOperator moves the selected rows with the ORCompact
 *     routing network (every row read and written regardless of the mask)
 *     instead of scatter + gather; with a Schema it finds the mask column by
 *     the reference's "$_mask" name, without one it takes the column index;
 *   - a Schema carries names and types only (ColumnInfo), the catalog's
 *     other duties are out of this path;
 *   - failures come back as Result::failure with the library's message;
 *     TaskManager methods throw ExecutionError, as CHECKED_CALL_THROW does
 *     (include/cuda/helper.hpp:42-48).
 */
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <functional>
#include <future>
#include <map>
#include <memory>
#include <mutex>
#include <optional>
#include <random>
#include <set>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <variant>
#include <vector>

#include "vo_api.h"

namespace velodb_amd {

class ExecutionError : public std::runtime_error {
public:
    explicit ExecutionError(const std::string &m) : std::runtime_error(m) {}
};

inline void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw ExecutionError(std::string(what) + ": " + hipGetErrorString(e));
}

// ----------------------------------------------------------------- Result

template <typename T, typename E = std::string>
class Result {
public:
    static Result success(T &&v) { return Result(std::in_place_index<0>, std::move(v)); }
    static Result failure(E e) { return Result(std::in_place_index<1>, std::move(e)); }
    explicit operator bool() const noexcept { return v_.index() == 0; }
    T &value() & { return std::get<0>(v_); }
    T &&value() && { return std::get<0>(std::move(v_)); }
    const E &error() const & { return std::get<1>(v_); }

private:
    template <size_t I, class A>
    Result(std::in_place_index_t<I> i, A &&a) : v_(i, std::forward<A>(a))
    {
    }
    std::variant<T, E> v_;
};

// ------------------------------------------------------------ TaskManager

class TaskManager {
public:
    explicit TaskManager(size_t queue_capacity = 1024, bool persistent = false, hipStream_t stream = nullptr)
    {
        vo_config cfg{};
        cfg.queue_capacity = static_cast<uint32_t>(queue_capacity);
        cfg.flags = persistent ? VO_CFG_PERSISTENT : 0;
        cfg.stream = stream;
        cfg.device = -1;
        const int st = vo_ctx_create(&cfg, &ctx_);
        if (st != VO_OK) throw ExecutionError(std::string("vo_ctx_create: ") + vo_status_string(st));
    }
    TaskManager(const TaskManager &) = delete;
    TaskManager &operator=(const TaskManager &) = delete;
    ~TaskManager()
    {
        if (ctx_) vo_ctx_destroy(ctx_);
    }

    bool start() { return is_running_ = (vo_queue_start(ctx_) == VO_OK); }
    bool stop(uint32_t timeout_ms = 5000)
    {
        const int st = vo_queue_stop(ctx_, timeout_ms);
        is_running_ = false;
        return st == VO_OK;
    }
    bool isRunning() const { return is_running_; }

    uint64_t submitCommand(vo_command &cmd)
    {
        const uint64_t seq = vo_submit(ctx_, &cmd);
        if (seq == 0) throw ExecutionError(vo_last_error(ctx_));
        return seq;
    }
    void waitCommand(uint64_t wait_id)
    {
        if (vo_wait(ctx_, wait_id, 0) != VO_OK) throw ExecutionError(vo_last_error(ctx_));
    }
    // throws ExecutionError on a non-zero status of a synchronous vo_* call
    void check(int status) const
    {
        if (status != VO_OK) throw ExecutionError(vo_last_error(ctx_));
    }
    vo_ctx *native() const { return ctx_; }

private:
    vo_ctx *ctx_ = nullptr;
    bool is_running_ = false;
};

class ExecutionContext {
public:
    explicit ExecutionContext(TaskManager &tm) : tm_(tm) {}
    TaskManager &getTaskManager() { return tm_; }

private:
    TaskManager &tm_;
};

// ----------------------------------------------------------------- Column

inline size_t type_width(uint32_t type_id)
{
    switch (type_id) {
    case VO_T_BOOLEAN: case VO_T_TINYINT: return 1;
    case VO_T_SMALLINT: return 2;
    case VO_T_INTEGER: case VO_T_FLOAT: case VO_T_DATE: return 4;
    case VO_T_BIGINT: case VO_T_DOUBLE: case VO_T_VARCHAR: case VO_T_UBIGINT: return 8;
    default: return 0;
    }
}

inline size_t next_pow2(size_t n, size_t floor_ = 4096)
{
    size_t p = floor_;
    while (p < n) p <<= 1;
    return p;
}

// A device column owning its buffers (data: capacity * width bytes; bitmap:
// capacity bytes or none).
class Column {
public:
    Column() = default;
    Column(uint32_t type_id, size_t capacity, bool nullable) : type_(type_id), capacity_(capacity)
    {
        if (type_width(type_id) == 0) throw ExecutionError("unsupported column type");
        hip_check(hipMalloc(&data_, capacity * type_width(type_id)), "hipMalloc");
        if (nullable) hip_check(hipMalloc(reinterpret_cast<void **>(&bitmap_), capacity), "hipMalloc");
    }
    // n values from host memory; valid (optional, n bytes) becomes the bitmap
    static Column fromHost(uint32_t type_id, const void *values, size_t n, const uint8_t *valid = nullptr)
    {
        Column c(type_id, next_pow2(n), valid != nullptr);
        hip_check(hipMemcpy(c.data_, values, n * type_width(type_id), hipMemcpyHostToDevice), "hipMemcpy");
        if (valid) hip_check(hipMemcpy(c.bitmap_, valid, n, hipMemcpyHostToDevice), "hipMemcpy");
        return c;
    }
    // the same through the context's padded staged transfer: the whole
    // capacity crosses the link (pad rows as zeros), so only the padded
    // size is visible on it (ValueVector::to, value_vector.hpp:140-294)
    static Column fromHost(TaskManager &tm, uint32_t type_id, const void *values, size_t n,
                           const uint8_t *valid = nullptr);
    template <typename T>
    std::vector<T> toHost(TaskManager &tm, size_t n) const;

    template <typename T>
    std::vector<T> toHost(size_t n) const
    {
        if (sizeof(T) != type_width(type_)) throw ExecutionError("toHost: element size does not match the column type");
        std::vector<T> v(n);
        if (n) hip_check(hipMemcpy(v.data(), data_, n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
        return v;
    }
    std::vector<uint8_t> bitmapToHost(size_t n) const
    {
        std::vector<uint8_t> v(n, 1);
        if (bitmap_ && n) hip_check(hipMemcpy(v.data(), bitmap_, n, hipMemcpyDeviceToHost), "hipMemcpy");
        return v;
    }

    Column(Column &&o) noexcept { *this = std::move(o); }
    Column &operator=(Column &&o) noexcept
    {
        if (this != &o) {
            release();
            type_ = o.type_;
            capacity_ = o.capacity_;
            data_ = std::exchange(o.data_, nullptr);
            bitmap_ = std::exchange(o.bitmap_, nullptr);
        }
        return *this;
    }
    Column(const Column &) = delete;
    Column &operator=(const Column &) = delete;
    ~Column() { release(); }

    uint32_t getTypeId() const { return type_; }
    size_t capacity() const { return capacity_; }
    void *rawData() { return data_; }
    const void *rawData() const { return data_; }
    uint8_t *rawBitmapData() { return bitmap_; }
    const uint8_t *bitmap() const { return bitmap_; }
    // new device buffers for a permuted copy (Column::getDeviceBuffer analogue)
    void *getDeviceBuffer() const
    {
        void *p = nullptr;
        hip_check(hipMalloc(&p, capacity_ * type_width(type_)), "hipMalloc");
        return p;
    }
    uint8_t *getDeviceBitmapBuffer() const
    {
        if (!bitmap_) return nullptr;
        void *p = nullptr;
        hip_check(hipMalloc(&p, capacity_), "hipMalloc");
        return static_cast<uint8_t *>(p);
    }
    // a column over buffers the caller allocated (createFromDeviceBuffers,
    // column.cpp:90-109): takes ownership
    static Column createFromDeviceBuffers(uint32_t type_id, void *data, uint8_t *bitmap, size_t capacity)
    {
        Column c;
        c.type_ = type_id;
        c.capacity_ = capacity;
        c.data_ = data;
        c.bitmap_ = bitmap;
        return c;
    }
    // take ownership of new buffers, freeing the old ones (setFromDeviceBuffers)
    void setFromDeviceBuffers(void *data, uint8_t *bitmap)
    {
        if (data) {
            (void)hipFree(data_);
            data_ = data;
        }
        if (bitmap) {
            (void)hipFree(bitmap_);
            bitmap_ = bitmap;
        }
    }

private:
    void release()
    {
        if (data_) (void)hipFree(data_);
        if (bitmap_) (void)hipFree(bitmap_);
        data_ = nullptr;
        bitmap_ = nullptr;
    }
    uint32_t type_ = 0;
    size_t capacity_ = 0;
    void *data_ = nullptr;
    uint8_t *bitmap_ = nullptr;
};

// Largest batch an operator hands downstream (include/common/constants.hpp.in:16).
constexpr size_t MAX_BATCH_SIZE = size_t(1) << 23;

class RowBatch {
public:
    RowBatch() = default;
    RowBatch(std::vector<Column> cols, size_t rows) : cols_(std::move(cols)), rows_(rows) {}
    size_t getRowCount() const { return rows_; }
    size_t getColumnCount() const { return cols_.size(); }
    Column &getColumn(size_t i) { return cols_.at(i); }
    const Column &getColumn(size_t i) const { return cols_.at(i); }
    std::vector<Column> &getColumns() { return cols_; }
    void setRowCount(size_t n) { rows_ = n; }

    // Rows [start, end) as a new batch with its own device columns of
    // power-of-two capacity (RowBatch::slice, src/catalog/row_batch.cpp:67-79).
    RowBatch slice(size_t start, size_t end) const
    {
        if (start > end || end > rows_) throw ExecutionError("Invalid slice range");
        const size_t n = end - start, cap = next_pow2(n);
        std::vector<Column> out;
        out.reserve(cols_.size());
        for (const Column &c : cols_) {
            const size_t w = type_width(c.getTypeId());
            Column o(c.getTypeId(), cap, c.bitmap() != nullptr);
            if (n) {
                hip_check(hipMemcpy(o.rawData(), static_cast<const char *>(c.rawData()) + start * w, n * w,
                                    hipMemcpyDeviceToDevice),
                          "hipMemcpy");
                if (c.bitmap())
                    hip_check(hipMemcpy(o.rawBitmapData(), c.bitmap() + start, n, hipMemcpyDeviceToDevice),
                              "hipMemcpy");
            }
            out.push_back(std::move(o));
        }
        return RowBatch(std::move(out), n);
    }

private:
    std::vector<Column> cols_;
    size_t rows_ = 0;
};

// ----------------------------------------------------------------- Schema

// A column's name and type (ColumnInfo, include/catalog/column.hpp:15-40).
class ColumnInfo {
public:
    ColumnInfo(std::string name, uint32_t type_id, bool is_nullable = true, bool is_unique = false,
               bool is_primary_key = false)
        : name_(std::move(name)), type_(type_id), nullable_(is_nullable), unique_(is_unique), pk_(is_primary_key)
    {
    }
    const std::string &getName() const { return name_; }
    uint32_t getTypeId() const { return type_; }
    bool isNullable() const { return nullable_; }
    bool isUnique() const { return unique_; }
    bool isPrimaryKey() const { return pk_; }

private:
    std::string name_;
    uint32_t type_;
    bool nullable_, unique_, pk_;
};

// An operator's output columns (Schema, include/catalog/schema.hpp:13-45).
class Schema {
public:
    Schema() = default;
    explicit Schema(std::vector<ColumnInfo> columns)
    {
        for (ColumnInfo &c : columns) addColumnInfo(std::move(c));
    }
    void addColumnInfo(ColumnInfo column)
    {
        index_[column.getName()] = cols_.size();
        cols_.push_back(std::move(column));
    }
    const ColumnInfo &getColumnInfo(size_t index) const { return cols_.at(index); }
    const ColumnInfo &getColumnInfo(const std::string &name) const { return cols_.at(getColumnIndex(name)); }
    size_t getColumnIndex(const std::string &name) const
    {
        auto it = index_.find(name);
        if (it == index_.end()) throw ExecutionError("Column not found: " + name);
        return it->second;
    }
    size_t getColumnCount() const { return cols_.size(); }
    bool hasColumn(const std::string &name) const { return index_.count(name) != 0; }
    Schema clone() const { return Schema(cols_); }

private:
    std::vector<ColumnInfo> cols_;
    std::unordered_map<std::string, size_t> index_;
};

class Table; // source tables of a join (catalog), passed through unused as in the reference

inline Column Column::fromHost(TaskManager &tm, uint32_t type_id, const void *values, size_t n, const uint8_t *valid)
{
    Column c(type_id, next_pow2(n), valid != nullptr);
    const size_t w = type_width(type_id);
    tm.check(vo_transfer_to_device(tm.native(), c.data_, values, n * w, c.capacity_ * w, nullptr));
    if (valid) tm.check(vo_transfer_to_device(tm.native(), c.bitmap_, valid, n, c.capacity_, nullptr));
    return c;
}

template <typename T>
std::vector<T> Column::toHost(TaskManager &tm, size_t n) const
{
    if (sizeof(T) != type_width(type_)) throw ExecutionError("toHost: element size does not match the column type");
    std::vector<T> v(n);
    tm.check(vo_transfer_to_host(tm.native(), v.data(), data_, n * sizeof(T), capacity_ * sizeof(T), nullptr));
    return v;
}

// --------------------------------------------------------------- operators

class AbstractOperator {
public:
    explicit AbstractOperator(ExecutionContext &context) : context_(context) {}
    AbstractOperator(ExecutionContext &context, Schema output_schema)
        : context_(context), output_schema_(std::move(output_schema))
    {
    }
    virtual ~AbstractOperator() = default;
    AbstractOperator(const AbstractOperator &) = delete;
    AbstractOperator &operator=(const AbstractOperator &) = delete;
    // an empty batch (no rows) ends the stream
    virtual Result<RowBatch> next() = 0;
    const Schema &getOutputSchema() const { return output_schema_; }

protected:
    // Every batch of the child, end to end, in one batch of power-of-two
    // capacity (AbstractOperator::collectBatches, src/operator/abstract_operator.cpp:17-81):
    // the first batch is kept when it is the only one; otherwise each column
    // is copied batch after batch into a new buffer.  A column that has a
    // null bitmap in any batch gets one, with the rows of the batches
    // without one valid.  A child failure throws ExecutionError, as the
    // reference does.
    RowBatch collectBatches(AbstractOperator &child)
    {
        std::vector<RowBatch> buffer;
        size_t n_rows = 0;
        for (;;) {
            auto r = child.next();
            if (!r) throw ExecutionError("Error fetching batch from child operator: " + r.error());
            RowBatch b = std::move(r).value();
            if (b.getRowCount() == 0) break;
            n_rows += b.getRowCount();
            buffer.push_back(std::move(b));
        }
        if (buffer.empty()) return RowBatch();
        if (buffer.size() == 1) return std::move(buffer.front());
        const size_t cap = next_pow2(n_rows), n_cols = buffer.front().getColumnCount();
        std::vector<Column> cols;
        cols.reserve(n_cols);
        for (size_t c = 0; c < n_cols; ++c) {
            const uint32_t type = buffer.front().getColumn(c).getTypeId();
            bool nullable = false;
            for (RowBatch &b : buffer) {
                if (b.getColumnCount() != n_cols || b.getColumn(c).getTypeId() != type)
                    throw ExecutionError("collectBatches: batches differ in their columns");
                nullable = nullable || b.getColumn(c).bitmap() != nullptr;
            }
            Column out(type, cap, nullable);
            const size_t w = type_width(type);
            size_t off = 0;
            for (RowBatch &b : buffer) {
                const Column &src = b.getColumn(c);
                const size_t rows = b.getRowCount();
                hip_check(hipMemcpyAsync(static_cast<char *>(out.rawData()) + off * w, src.rawData(), rows * w,
                                         hipMemcpyDeviceToDevice, nullptr),
                          "hipMemcpyAsync");
                if (nullable) {
                    if (src.bitmap())
                        hip_check(hipMemcpyAsync(out.rawBitmapData() + off, src.bitmap(), rows,
                                                 hipMemcpyDeviceToDevice, nullptr),
                                  "hipMemcpyAsync");
                    else
                        hip_check(hipMemsetAsync(out.rawBitmapData() + off, 1, rows, nullptr), "hipMemsetAsync");
                }
                off += rows;
            }
            cols.push_back(std::move(out));
        }
        hip_check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
        return RowBatch(std::move(cols), n_rows);
    }

    ExecutionContext &context_;
    Schema output_schema_;
};

// Source of prepared batches, one per next() (stands in for a scan of
// device columns).
class BatchSourceOperator : public AbstractOperator {
public:
    BatchSourceOperator(ExecutionContext &context, RowBatch batch) : AbstractOperator(context)
    {
        batches_.push_back(std::move(batch));
    }
    BatchSourceOperator(ExecutionContext &context, std::vector<RowBatch> batches)
        : AbstractOperator(context), batches_(std::move(batches))
    {
    }
    Result<RowBatch> next() override
    {
        if (at_ >= batches_.size()) return Result<RowBatch>::success(RowBatch());
        return Result<RowBatch>::success(std::move(batches_[at_++]));
    }

private:
    std::vector<RowBatch> batches_;
    size_t at_ = 0;
};

// Oblivious multi-column sort: every child batch gathered into one
// (collectBatches), OP_SORT over the key columns (pads written in place,
// int64 row indices produced), then OP_PERMUTE of every other column and of
// every null bitmap (sort_operator.cpp:21-134).  The sorted rows come back
// in slices of at most MAX_BATCH_SIZE rows, one per next(), then an empty
// batch (sort_operator.cpp:24-33, 127-133).
class SortOperator : public AbstractOperator {
public:
    SortOperator(ExecutionContext &context, Schema output_schema, std::unique_ptr<AbstractOperator> child,
                 std::vector<size_t> order_indices, std::vector<bool> ascending_flags)
        : AbstractOperator(context, std::move(output_schema)), child_(std::move(child)),
          order_(std::move(order_indices)), asc_(std::move(ascending_flags))
    {
    }
    SortOperator(ExecutionContext &context, std::unique_ptr<AbstractOperator> child, std::vector<size_t> order_indices,
                 std::vector<bool> ascending_flags)
        : SortOperator(context, Schema(), std::move(child), std::move(order_indices), std::move(ascending_flags))
    {
    }

    // rows per output slice (MAX_BATCH_SIZE unless set smaller, e.g. by tests)
    void setMaxBatchSize(size_t rows) { max_batch_ = std::max<size_t>(1, rows); }

    Result<RowBatch> next() override
    {
        if (!sorted_) {
            try {
                RowBatch gathered = collectBatches(*child_);
                sorted_ = true;
                if (gathered.getRowCount() == 0) return Result<RowBatch>::success(RowBatch());
                run(gathered);
                sorted_batch_ = std::move(gathered);
                offset_ = 0;
            } catch (const std::exception &e) {
                sorted_ = true;
                return Result<RowBatch>::failure(e.what());
            }
        }
        const size_t total = sorted_batch_.getRowCount();
        if (offset_ >= total) return Result<RowBatch>::success(RowBatch());
        if (offset_ == 0 && total <= max_batch_) { // one slice: the batch itself
            offset_ = total;
            return Result<RowBatch>::success(std::move(sorted_batch_));
        }
        const size_t end = std::min(offset_ + max_batch_, total);
        try {
            RowBatch slice = sorted_batch_.slice(offset_, end);
            offset_ = end;
            return Result<RowBatch>::success(std::move(slice));
        } catch (const std::exception &e) {
            return Result<RowBatch>::failure(e.what());
        }
    }

private:
    void run(RowBatch &batch)
    {
        const size_t n = batch.getRowCount();
        if (n == 0) return;
        if (order_.empty() || order_.size() > VO_MAX_SORT_COLUMNS || asc_.size() != order_.size())
            throw ExecutionError("SortOperator: 1..8 sort columns, one ascending flag each");
        const size_t n_pad = batch.getColumn(order_[0]).capacity();
        TaskManager &tm = context_.getTaskManager();
        int64_t *d_indices = nullptr;
        hip_check(hipMalloc(reinterpret_cast<void **>(&d_indices), n_pad * sizeof(int64_t)), "hipMalloc");
        std::unique_ptr<int64_t, void (*)(int64_t *)> guard(d_indices, [](int64_t *p) { (void)hipFree(p); });
        std::vector<bool> is_key(batch.getColumnCount(), false);
        vo_command sort_cmd{};
        sort_cmd.opcode = VO_OP_SORT;
        sort_cmd.args.sort.indices = d_indices;
        sort_cmd.args.sort.n_sort_columns = order_.size();
        sort_cmd.args.sort.n_rows = n;
        sort_cmd.args.sort.n_padded_rows = n_pad;
        for (size_t i = 0; i < order_.size(); ++i) {
            Column &c = batch.getColumn(order_[i]);
            if (c.capacity() != n_pad) throw ExecutionError("SortOperator: key columns differ in capacity");
            is_key[order_[i]] = true;
            sort_cmd.args.sort.sort_cols[i] = c.rawData();
            sort_cmd.args.sort.col_types[i] = c.getTypeId();
            sort_cmd.args.sort.ascending_flags[i] = asc_[i] ? 1 : 0;
        }
        tm.waitCommand(tm.submitCommand(sort_cmd));
        for (size_t i = 0; i < batch.getColumnCount(); ++i) {
            Column &c = batch.getColumn(i);
            void *data_buf = nullptr;
            uint64_t last = 0;
            if (!is_key[i]) {
                data_buf = c.getDeviceBuffer();
                vo_command p{};
                p.opcode = VO_OP_PERMUTE;
                p.args.permute = {data_buf, c.rawData(), d_indices, n, c.getTypeId()};
                last = tm.submitCommand(p);
            }
            uint8_t *bits_buf = c.getDeviceBitmapBuffer();
            if (bits_buf) {
                vo_command p{};
                p.opcode = VO_OP_PERMUTE;
                p.args.permute = {bits_buf, c.rawBitmapData(), d_indices, n, VO_T_BOOLEAN};
                last = tm.submitCommand(p);
            }
            if (last) tm.waitCommand(last);
            c.setFromDeviceBuffers(data_buf, bits_buf);
        }
    }

    std::unique_ptr<AbstractOperator> child_;
    std::vector<size_t> order_;
    std::vector<bool> asc_;
    bool sorted_ = false;
    RowBatch sorted_batch_; // the whole sorted result, handed out in slices
    size_t offset_ = 0;
    size_t max_batch_ = MAX_BATCH_SIZE;
};

// Predicate terms evaluated on the device (OP_PREDICATE_MASK): a column
// compared with a constant in the column's type, as ComparisonExpression
// does with a scalar (src/expression/comparison_expression.cpp:13-92); the
// terms of one predicate are joined by one ConnectiveType
// (src/expression/logical_expression.cpp:61-150).  A column's NULL rows
// (bitmap byte 0) make the row unselected.  The enums keep the reference's
// order (include/expression/expression.hpp:34-43, without LIKE).
enum class ComparisonType { EQUAL, NOT_EQUAL, LESS_THAN, LESS_THAN_OR_EQUAL, GREATER_THAN, GREATER_THAN_OR_EQUAL };
enum class ConnectiveType { AND, OR };

struct PredicateTerm {
    size_t column;       // column index in the batch
    ComparisonType cmp;
    uint64_t constant;   // the constant's bit pattern in the column's type

    // make(column, cmp, value): value of the column's C type (int32_t for
    // INTEGER, double for DOUBLE, ...)
    template <typename T>
    static PredicateTerm make(size_t column, ComparisonType cmp, T value)
    {
        static_assert(sizeof(T) <= 8, "scalar constant");
        uint64_t bits = 0;
        std::memcpy(&bits, &value, sizeof(T));
        return PredicateTerm{column, cmp, bits};
    }
};

// Oblivious filter compaction: the selected rows move to the front, in
// order, in every column marked in compact_columns (data and null bitmap);
// the batch's row count becomes the selected count.  Columns not marked keep
// their rows, as the reference leaves them
// (filter_compaction_operator.cpp:112-143).  The selection is either a mask
// column (BOOLEAN, non-zero = selected; the scan's "$_mask") or a predicate
// evaluated on the device, which the reference evaluates on the host
// (seq_scan_operator.cpp:38-45).
// Batches stream through one at a time; while one is compacted on the
// device, the next is fetched from the child on another thread, as the
// reference's prefetch does (filter_compaction_operator.cpp:38-62, 186-233).
class FilterCompactionOperator : public AbstractOperator {
public:
    // The reference's form: the mask column is the schema's "$_mask" (or a
    // name ending in ".$_mask"); compact_columns defaults to every column and
    // must match the schema (filter_compaction_operator.cpp:16-28, 69-82).
    FilterCompactionOperator(ExecutionContext &context, Schema output_schema, std::unique_ptr<AbstractOperator> child,
                             std::vector<bool> compact_columns = {})
        : AbstractOperator(context, std::move(output_schema)), child_(std::move(child)), mask_index_(kNoMask),
          compact_(std::move(compact_columns))
    {
        const size_t nc = output_schema_.getColumnCount();
        for (size_t i = 0; i < nc && mask_index_ == kNoMask; ++i) {
            const std::string &name = output_schema_.getColumnInfo(i).getName();
            if (name == "$_mask" || (name.size() > 7 && name.compare(name.size() - 7, 7, ".$_mask") == 0))
                mask_index_ = i;
        }
        if (mask_index_ == kNoMask) mask_index_ = output_schema_.getColumnIndex("$_mask"); // throws, as the reference
        if (compact_.empty()) compact_.assign(nc, true);
        if (compact_.size() != nc) throw ExecutionError("Compaction mask must match the output schema");
    }

    FilterCompactionOperator(ExecutionContext &context, std::unique_ptr<AbstractOperator> child, size_t mask_index,
                             std::vector<bool> compact_columns = {})
        : AbstractOperator(context), child_(std::move(child)), mask_index_(mask_index),
          compact_(std::move(compact_columns))
    {
    }

    FilterCompactionOperator(ExecutionContext &context, std::unique_ptr<AbstractOperator> child,
                             std::vector<PredicateTerm> predicate, ConnectiveType connective = ConnectiveType::AND,
                             std::vector<bool> compact_columns = {})
        : AbstractOperator(context), child_(std::move(child)), mask_index_(kNoMask), pred_(std::move(predicate)),
          conn_(connective), compact_(std::move(compact_columns))
    {
    }

    ~FilterCompactionOperator() override
    {
        if (prefetch_.valid()) prefetch_.wait();
    }

    Result<RowBatch> next() override
    {
        Result<RowBatch> in = Result<RowBatch>::success(RowBatch());
        if (first_call_) {
            first_call_ = false;
            in = child_->next();
        } else {
            if (!prefetch_.valid()) return Result<RowBatch>::success(RowBatch()); // the stream ended
            in = prefetch_.get();
        }
        if (!in) return in;
        RowBatch batch = std::move(in).value();
        if (batch.getRowCount() == 0) return Result<RowBatch>::success(std::move(batch));
        // the next batch is fetched while this one is compacted
        prefetch_ = std::async(std::launch::async, [this] { return child_->next(); });
        try {
            run(batch);
        } catch (const std::exception &e) {
            return Result<RowBatch>::failure(e.what());
        }
        return Result<RowBatch>::success(std::move(batch));
    }

private:
    static constexpr size_t kNoMask = ~(size_t)0;

    // the predicate's mask, one byte per padded row, in a new device buffer
    uint8_t *evaluate(RowBatch &batch, size_t n_pad)
    {
        if (pred_.empty() || pred_.size() > VO_MAX_PREDICATE_TERMS)
            throw ExecutionError("FilterCompactionOperator: 1..8 predicate terms");
        std::vector<vo_predicate_term> terms;
        for (const PredicateTerm &t : pred_) {
            if (t.column >= batch.getColumnCount())
                throw ExecutionError("FilterCompactionOperator: predicate column out of range");
            const Column &c = batch.getColumn(t.column);
            terms.push_back(vo_predicate_term{c.rawData(), c.bitmap(), t.constant, c.getTypeId(),
                                              static_cast<uint32_t>(t.cmp)});
        }
        uint8_t *mask = nullptr;
        hip_check(hipMalloc(reinterpret_cast<void **>(&mask), n_pad), "hipMalloc");
        vo_predicate_args a{};
        a.terms = terms.data();
        a.n_terms = (uint32_t)terms.size();
        a.connective = conn_ == ConnectiveType::AND ? VO_CONN_AND : VO_CONN_OR;
        a.out_mask = mask;
        a.n = batch.getRowCount();
        a.n_pad = n_pad;
        TaskManager &tm = context_.getTaskManager();
        const int st = vo_predicate_mask(tm.native(), &a);
        if (st != VO_OK) {
            (void)hipFree(mask);
            tm.check(st);
        }
        return mask;
    }

    void run(RowBatch &batch)
    {
        const size_t n = batch.getRowCount();
        const uint8_t *mask = nullptr;
        size_t n_pad = 0;
        std::unique_ptr<uint8_t, void (*)(uint8_t *)> pred_mask(nullptr, [](uint8_t *p) { (void)hipFree(p); });
        if (mask_index_ != kNoMask) {
            if (mask_index_ >= batch.getColumnCount())
                throw ExecutionError("FilterCompactionOperator: mask column out of range");
            Column &m = batch.getColumn(mask_index_);
            if (type_width(m.getTypeId()) != 1) throw ExecutionError("FilterCompactionOperator: mask must be 1 byte/row");
            mask = static_cast<const uint8_t *>(m.rawData());
            n_pad = m.capacity();
        } else {
            n_pad = batch.getColumn(0).capacity();
            pred_mask.reset(evaluate(batch, n_pad));
            mask = pred_mask.get();
        }
        std::vector<std::pair<void *, uint32_t>> cols; // every buffer moved by the network
        for (size_t i = 0; i < batch.getColumnCount(); ++i) {
            if (i == mask_index_) continue;
            if (!compact_.empty() && (i >= compact_.size() || !compact_[i])) continue;
            Column &c = batch.getColumn(i);
            if (c.capacity() != n_pad) throw ExecutionError("FilterCompactionOperator: columns differ in capacity");
            cols.emplace_back(c.rawData(), (uint32_t)type_width(c.getTypeId()));
            if (c.rawBitmapData()) cols.emplace_back(c.rawBitmapData(), 1u);
        }
        uint64_t *d_count = nullptr;
        hip_check(hipMalloc(reinterpret_cast<void **>(&d_count), sizeof(uint64_t)), "hipMalloc");
        std::unique_ptr<uint64_t, void (*)(uint64_t *)> guard(d_count, [](uint64_t *p) { (void)hipFree(p); });
        TaskManager &tm = context_.getTaskManager();
        // the routing depends on the mask only: each group of up to 8
        // buffers is moved by the same network
        for (size_t at = 0; at < cols.size(); at += VO_MAX_SORT_COLUMNS) {
            vo_compact_args a{};
            a.mask = mask;
            const size_t k = std::min<size_t>(VO_MAX_SORT_COLUMNS, cols.size() - at);
            for (size_t j = 0; j < k; ++j) {
                a.cols[j] = cols[at + j].first;
                a.col_bytes[j] = cols[at + j].second;
            }
            a.n_cols = (uint32_t)k;
            a.n = n;
            a.n_pad = n_pad;
            a.out_count = d_count;
            tm.check(vo_compact(tm.native(), &a));
        }
        if (cols.empty()) {
            // nothing to move: the count alone (a compaction of a mask copy)
            void *tmp = nullptr;
            hip_check(hipMalloc(&tmp, n_pad), "hipMalloc");
            std::unique_ptr<void, void (*)(void *)> g2(tmp, [](void *p) { (void)hipFree(p); });
            hip_check(hipMemcpy(tmp, mask, n_pad, hipMemcpyDeviceToDevice), "hipMemcpy");
            vo_compact_args a{};
            a.mask = mask;
            a.cols[0] = tmp;
            a.col_bytes[0] = 1;
            a.n_cols = 1;
            a.n = n;
            a.n_pad = n_pad;
            a.out_count = d_count;
            tm.check(vo_compact(tm.native(), &a));
        }
        uint64_t count = 0;
        hip_check(hipMemcpy(&count, d_count, sizeof(count), hipMemcpyDeviceToHost), "hipMemcpy");
        batch.setRowCount(count);
    }

    std::unique_ptr<AbstractOperator> child_;
    size_t mask_index_;
    std::vector<PredicateTerm> pred_;
    ConnectiveType conn_ = ConnectiveType::AND;
    std::vector<bool> compact_;
    bool first_call_ = true;
    std::future<Result<RowBatch>> prefetch_; // the child's next batch, fetched on another thread
};

enum class JoinType { INNER, LEFT, RIGHT, FULL };

// Equi-join of two inputs on one key column each (SortMergeJoinOperator,
// sort_merge_join_operator.cpp:27-257): the pairs of row indices with equal
// keys come from the oblivious OP_SORT_MERGE_JOIN_COUNT / _WRITE (every
// input row read and written by every pass; the pair count is the one value
// that leaves the device, as in the reference), then every column and null
// bitmap of both inputs is gathered by OP_PERMUTE, left columns first, as
// :151-235 does.  The pair computation is oblivious; the permute reads each
// input at the pair indices, which depend on the data, exactly as the
// reference's does.  The inputs need not be
// sorted (the reference requires both sorted by key, :58-59).  The output
// rows come ordered by (key, left row, right row) with a power-of-two
// capacity; one batch, then an empty one.  Every batch of each child is
// gathered first (collectBatches, :37-49).  INNER joins only, as in the
// reference.
class SortMergeJoinOperator : public AbstractOperator {
public:
    // The reference's form (include/operator/sort_merge_join_operator.hpp:16-23);
    // the source tables are kept and not used, as there.
    SortMergeJoinOperator(ExecutionContext &context, Schema output_schema,
                          std::unique_ptr<AbstractOperator> left_child, std::unique_ptr<AbstractOperator> right_child,
                          std::pair<size_t, size_t> join_key_indices, std::vector<const Table *> left_source_tables,
                          std::vector<const Table *> right_source_tables, JoinType join_type = JoinType::INNER,
                          bool oblivious_materialization = false)
        : AbstractOperator(context, std::move(output_schema)), left_(std::move(left_child)),
          right_(std::move(right_child)), keys_(join_key_indices), left_tables_(std::move(left_source_tables)),
          right_tables_(std::move(right_source_tables)), type_(join_type), oblivious_(oblivious_materialization)
    {
    }
    // oblivious_materialization: gather the columns with OP_GATHER_OBLIVIOUS
    // (no address depends on the pairs) instead of the reference's permute
    SortMergeJoinOperator(ExecutionContext &context, std::unique_ptr<AbstractOperator> left_child,
                          std::unique_ptr<AbstractOperator> right_child, std::pair<size_t, size_t> join_key_indices,
                          JoinType join_type = JoinType::INNER, bool oblivious_materialization = false)
        : SortMergeJoinOperator(context, Schema(), std::move(left_child), std::move(right_child), join_key_indices, {},
                                {}, join_type, oblivious_materialization)
    {
    }

    Result<RowBatch> next() override
    {
        if (joined_) return Result<RowBatch>::success(RowBatch());
        joined_ = true;
        if (type_ != JoinType::INNER) return Result<RowBatch>::failure("SortMergeJoinOperator supports INNER joins");
        RowBatch lb, rb;
        try {
            lb = collectBatches(*left_);
            if (lb.getRowCount() == 0) return Result<RowBatch>::success(RowBatch());
            rb = collectBatches(*right_);
            if (rb.getRowCount() == 0) return Result<RowBatch>::success(RowBatch());
        } catch (const std::exception &e) {
            return Result<RowBatch>::failure(e.what());
        }
        if (keys_.first >= lb.getColumnCount()) return Result<RowBatch>::failure("Left join key index out of bounds");
        if (keys_.second >= rb.getColumnCount())
            return Result<RowBatch>::failure("Right join key index out of bounds");
        try {
            return Result<RowBatch>::success(run(lb, rb));
        } catch (const std::exception &e) {
            return Result<RowBatch>::failure(e.what());
        }
    }

private:
    RowBatch run(RowBatch &lb, RowBatch &rb)
    {
        Column &lk = lb.getColumn(keys_.first), &rk = rb.getColumn(keys_.second);
        if (lk.getTypeId() != rk.getTypeId()) throw ExecutionError("SortMergeJoinOperator: key types differ");
        TaskManager &tm = context_.getTaskManager();
        uint64_t *d_count = nullptr;
        hip_check(hipMalloc(reinterpret_cast<void **>(&d_count), sizeof(uint64_t)), "hipMalloc");
        std::unique_ptr<uint64_t, void (*)(uint64_t *)> guard(d_count, [](uint64_t *p) { (void)hipFree(p); });
        vo_join_pairs_args a{};
        a.left_keys = lk.rawData();
        a.n_left = lb.getRowCount();
        a.right_keys = rk.rawData();
        a.n_right = rb.getRowCount();
        a.type_id = lk.getTypeId();
        a.out_count = d_count;
        // the COUNT -> WRITE routing state in a buffer of this operator, as
        // the reference's MergeJoinBlock buffers (commands.hpp:82-98): joins
        // of other operators on the same TaskManager may run in between
        void *state = nullptr;
        hip_check(hipMalloc(&state, vo_join_pairs_state_bytes(a.n_left, a.n_right)), "hipMalloc");
        std::unique_ptr<void, void (*)(void *)> state_guard(state, [](void *p) { (void)hipFree(p); });
        a.state = state;
        vo_command count{};
        count.opcode = VO_OP_SORT_MERGE_JOIN_COUNT;
        count.args.join_pairs = a;
        tm.waitCommand(tm.submitCommand(count));
        uint64_t n_out = 0;
        hip_check(hipMemcpy(&n_out, d_count, sizeof(n_out), hipMemcpyDeviceToHost), "hipMemcpy");
        if (n_out == 0) return RowBatch();
        const size_t cap = next_pow2(n_out);
        Column li = Column(VO_T_BIGINT, cap, false), ri = Column(VO_T_BIGINT, cap, false);
        a.out_count = nullptr;
        a.n_out = n_out;
        a.out_left = static_cast<int64_t *>(li.rawData());
        a.out_right = static_cast<int64_t *>(ri.rawData());
        a.out_cap = cap;
        vo_command write{};
        write.opcode = VO_OP_SORT_MERGE_JOIN_WRITE;
        write.args.join_pairs = a;
        tm.waitCommand(tm.submitCommand(write));
        std::vector<Column> out;
        out.reserve(lb.getColumnCount() + rb.getColumnCount());
        uint64_t last = 0;
        auto gather = [&](RowBatch &b, const Column &idx) {
            for (size_t i = 0; i < b.getColumnCount(); ++i) {
                Column &c = b.getColumn(i);
                void *data = nullptr;
                hip_check(hipMalloc(&data, cap * type_width(c.getTypeId())), "hipMalloc");
                uint8_t *bits = nullptr;
                if (c.rawBitmapData()) hip_check(hipMalloc(reinterpret_cast<void **>(&bits), cap), "hipMalloc");
                out.push_back(Column::createFromDeviceBuffers(c.getTypeId(), data, bits, cap));
                const int64_t *ids = static_cast<const int64_t *>(idx.rawData());
                if (oblivious_) {
                    last = submit_gather(tm, c.rawData(), b.getRowCount(), ids, n_out, data,
                                         (uint32_t)type_width(c.getTypeId()));
                    if (bits) last = submit_gather(tm, c.rawBitmapData(), b.getRowCount(), ids, n_out, bits, 1);
                    continue;
                }
                vo_command p{};
                p.opcode = VO_OP_PERMUTE;
                p.args.permute = {data, c.rawData(), ids, n_out, c.getTypeId()};
                last = tm.submitCommand(p);
                if (bits) {
                    vo_command pb{};
                    pb.opcode = VO_OP_PERMUTE;
                    pb.args.permute = {bits, c.rawBitmapData(), ids, n_out, VO_T_BOOLEAN};
                    last = tm.submitCommand(pb);
                }
            }
        };
        gather(lb, li);
        gather(rb, ri);
        if (last) tm.waitCommand(last);
        return RowBatch(std::move(out), n_out);
    }

    static uint64_t submit_gather(TaskManager &tm, const void *table, size_t n_table, const int64_t *ids, size_t n,
                                  void *out, uint32_t width)
    {
        vo_command g{};
        g.opcode = VO_OP_GATHER_OBLIVIOUS;
        g.args.gather_oblivious = {table, n_table, ids, n, out, nullptr, width, 0};
        return tm.submitCommand(g);
    }

    std::unique_ptr<AbstractOperator> left_, right_;
    std::pair<size_t, size_t> keys_;
    std::vector<const Table *> left_tables_, right_tables_;
    JoinType type_;
    bool oblivious_;
    bool joined_ = false;
};

// Oblivious materialization of columns by record id
// (MaterializationOperator, src/operator/materialization_operator.cpp:362-517):
// for each requested column, out[i] = source[rowid[i]] through
// OP_GATHER_OBLIVIOUS, which sorts the table rows and the requests together
// instead of reading rows at positions (the reference's host bitonic sort,
// cached reads with dummy reads, restore).  The child yields the row-id
// columns (BIGINT); each output column names a row-id column of the child and
// a column of a source batch.  Null bitmaps are gathered with their values;
// an id outside the source gives 0 and, when the source column is nullable,
// a null.
struct MaterializedColumn {
    size_t rowid_index;     // BIGINT column of the child
    const RowBatch *source; // source table
    size_t column_index;    // column of the source
};

class MaterializationOperator : public AbstractOperator {
public:
    MaterializationOperator(ExecutionContext &context, std::unique_ptr<AbstractOperator> child,
                            std::vector<MaterializedColumn> columns)
        : AbstractOperator(context), child_(std::move(child)), cols_(std::move(columns))
    {
    }

    Result<RowBatch> next() override
    {
        auto in = child_->next();
        if (!in) return in;
        RowBatch ids = std::move(in).value();
        if (ids.getRowCount() == 0) return Result<RowBatch>::success(RowBatch());
        try {
            return Result<RowBatch>::success(run(ids));
        } catch (const std::exception &e) {
            return Result<RowBatch>::failure(e.what());
        }
    }

private:
    RowBatch run(RowBatch &ids)
    {
        const size_t n = ids.getRowCount(), cap = next_pow2(n);
        TaskManager &tm = context_.getTaskManager();
        std::vector<Column> out;
        uint64_t last = 0;
        for (const MaterializedColumn &m : cols_) {
            if (!m.source || m.rowid_index >= ids.getColumnCount() || m.column_index >= m.source->getColumnCount())
                throw ExecutionError("MaterializationOperator: column out of range");
            const Column &rid = ids.getColumn(m.rowid_index);
            if (rid.getTypeId() != VO_T_BIGINT) throw ExecutionError("MaterializationOperator: row ids must be BIGINT");
            const Column &src = m.source->getColumn(m.column_index);
            const size_t nt = m.source->getRowCount();
            Column o(src.getTypeId(), cap, src.bitmap() != nullptr);
            vo_command g{};
            g.opcode = VO_OP_GATHER_OBLIVIOUS;
            g.args.gather_oblivious = {src.rawData(), nt, static_cast<const int64_t *>(rid.rawData()), n, o.rawData(),
                                       nullptr, (uint32_t)type_width(src.getTypeId()), 0};
            last = tm.submitCommand(g);
            if (src.bitmap()) {
                // an id outside the source gathers 0: a null
                vo_command gb{};
                gb.opcode = VO_OP_GATHER_OBLIVIOUS;
                gb.args.gather_oblivious = {src.bitmap(), nt, static_cast<const int64_t *>(rid.rawData()), n,
                                            o.rawBitmapData(), nullptr, 1, 0};
                last = tm.submitCommand(gb);
            }
            out.push_back(std::move(o));
        }
        if (last) tm.waitCommand(last);
        return RowBatch(std::move(out), n);
    }

    std::unique_ptr<AbstractOperator> child_;
    std::vector<MaterializedColumn> cols_;
};

// Oblivious PK/FK join + SUM(value) GROUP BY group: dimension batch (key
// INTEGER, group INTEGER), fact batch (foreign key INTEGER, value BIGINT).
// Output: one batch of G rows, columns (group INTEGER, sum BIGINT), groups
// ascending.  dense_groups: group ids lie in [0, max_groups) (max_groups <=
// 64), a public bound of the schema.
class PKFKJoinGroupSumOperator : public AbstractOperator {
public:
    PKFKJoinGroupSumOperator(ExecutionContext &context, std::unique_ptr<AbstractOperator> dim,
                             std::unique_ptr<AbstractOperator> fact, size_t dim_key, size_t dim_group, size_t fact_fk,
                             size_t fact_val, size_t max_groups, bool dense_groups)
        : AbstractOperator(context), dim_(std::move(dim)), fact_(std::move(fact)), dk_(dim_key), dg_(dim_group),
          fk_(fact_fk), fv_(fact_val), max_groups_(max_groups), dense_(dense_groups)
    {
    }

    Result<RowBatch> next() override
    {
        if (done_) return Result<RowBatch>::success(RowBatch());
        done_ = true;
        try {
            // every batch of each side, as the sort-merge join gathers them
            RowBatch db = collectBatches(*dim_);
            RowBatch fb = collectBatches(*fact_);
            if (db.getRowCount() == 0 || fb.getRowCount() == 0) return Result<RowBatch>::success(RowBatch());
            return Result<RowBatch>::success(run(db, fb));
        } catch (const std::exception &e) {
            return Result<RowBatch>::failure(e.what());
        }
    }

private:
    RowBatch run(RowBatch &db, RowBatch &fb)
    {
        Column &dk = db.getColumn(dk_), &dg = db.getColumn(dg_), &fk = fb.getColumn(fk_), &fv = fb.getColumn(fv_);
        if (dk.getTypeId() != VO_T_INTEGER || dg.getTypeId() != VO_T_INTEGER || fk.getTypeId() != VO_T_INTEGER ||
            fv.getTypeId() != VO_T_BIGINT)
            throw ExecutionError("PKFKJoinGroupSumOperator: keys and groups INTEGER, values BIGINT");
        Column og(VO_T_INTEGER, next_pow2(max_groups_, 1), false), os(VO_T_BIGINT, next_pow2(max_groups_, 1), false);
        uint64_t *d_n = nullptr;
        hip_check(hipMalloc(reinterpret_cast<void **>(&d_n), sizeof(uint64_t)), "hipMalloc");
        std::unique_ptr<uint64_t, void (*)(uint64_t *)> guard(d_n, [](uint64_t *p) { (void)hipFree(p); });
        vo_join_args a{};
        a.dim_key = static_cast<const int32_t *>(dk.rawData());
        a.dim_group = static_cast<const int32_t *>(dg.rawData());
        a.n_dim = db.getRowCount();
        a.fact_fk = static_cast<const int32_t *>(fk.rawData());
        a.fact_val = static_cast<const int64_t *>(fv.rawData());
        a.n_fact = fb.getRowCount();
        a.n_pad = next_pow2(a.n_dim + a.n_fact);
        a.out_group = static_cast<int32_t *>(og.rawData());
        a.out_sum = static_cast<int64_t *>(os.rawData());
        a.out_n_groups = d_n;
        a.max_groups = max_groups_;
        a.dense_groups = dense_ ? 1 : 0;
        TaskManager &tm = context_.getTaskManager();
        tm.check(vo_pkfk_join_groupsum(tm.native(), &a));
        uint64_t g = 0;
        hip_check(hipMemcpy(&g, d_n, sizeof(g), hipMemcpyDeviceToHost), "hipMemcpy");
        if (g == ~0ull) throw ExecutionError("PKFKJoinGroupSumOperator: a fact value above INT64_MAX - 65");
        std::vector<Column> cols;
        cols.push_back(std::move(og));
        cols.push_back(std::move(os));
        return RowBatch(std::move(cols), std::min<uint64_t>(g, max_groups_));
    }

    std::unique_ptr<AbstractOperator> dim_, fact_;
    size_t dk_, dg_, fk_, fv_, max_groups_;
    bool dense_;
    bool done_ = false;
};

// ------------------------------------------------------------- plan nodes

// Plan nodes of this path: each builds its operator, and its children's,
// from createOperator(context), as the reference's planner nodes do
// (include/planner/abstract_plan_node.hpp:19-59).  The PlanType values are
// the reference's (:19-33), so code that switches on them keeps working.
// Without a Schema, a node names columns by index; toString() prints "#i".
// createOperator may be called more than once (the nodes keep their
// arguments and copy them into each operator).
enum class PlanType : uint16_t {
    INVALID = 0,
    SEQ_SCAN = 0b0000'0001,
    COMPACTION = 0b0000'0010,
    JOIN = 0b0000'1000,
    SORT_MERGE_JOIN = 0b0011'1000,
    MATERIALIZATION = 0b0100'0000,
    SORT = 0b1000'0000,
    AGGREGATE = 0b10'1000'0000,
};

class AbstractPlanNode {
public:
    explicit AbstractPlanNode(PlanType type) : type_(type) {}
    virtual ~AbstractPlanNode() = default;
    AbstractPlanNode(const AbstractPlanNode &) = delete;
    AbstractPlanNode &operator=(const AbstractPlanNode &) = delete;

    PlanType getPlanType() const { return type_; }
    // the output columns (AbstractPlanNode::getOutputSchema); empty when the
    // node names columns by index only
    const Schema &getOutputSchema() const { return output_schema_; }
    void setOutputSchema(Schema schema) { output_schema_ = std::move(schema); }
    const std::vector<std::unique_ptr<AbstractPlanNode>> &getChildren() const { return children_; }
    void addChild(std::unique_ptr<AbstractPlanNode> child) { children_.push_back(std::move(child)); }

    virtual std::unique_ptr<AbstractOperator> createOperator(ExecutionContext &context) const = 0;
    virtual std::string toString() const = 0;

protected:
    // children_[i]->createOperator(context), after checking the child count
    // (the reference asserts it, e.g. sort_plan_node.cpp:19)
    std::unique_ptr<AbstractOperator> child_op(ExecutionContext &context, size_t i, size_t expect,
                                               const char *node) const
    {
        if (children_.size() != expect)
            throw ExecutionError(std::string(node) + " must have exactly " + std::to_string(expect) +
                                 (expect == 1 ? " child" : " children"));
        return children_[i]->createOperator(context);
    }

    PlanType type_;
    Schema output_schema_;
    std::vector<std::unique_ptr<AbstractPlanNode>> children_;
};

// Leaf that produces one batch per operator it creates; `scan` stands in for
// the sequential scan of device columns (seq_scan_plan_node.cpp).
class ScanPlanNode : public AbstractPlanNode {
public:
    explicit ScanPlanNode(std::function<RowBatch()> scan, std::string name = "batch")
        : AbstractPlanNode(PlanType::SEQ_SCAN), scan_(std::move(scan)), name_(std::move(name))
    {
    }
    std::unique_ptr<AbstractOperator> createOperator(ExecutionContext &context) const override
    {
        return std::make_unique<BatchSourceOperator>(context, scan_());
    }
    std::string toString() const override { return "Scan(" + name_ + ")"; }

private:
    std::function<RowBatch()> scan_;
    std::string name_;
};

// SortPlanNode (include/planner/sort_plan_node.hpp, src/planner/sort_plan_node.cpp:17-27)
class SortPlanNode : public AbstractPlanNode {
public:
    SortPlanNode(std::vector<size_t> order_indices, std::vector<bool> ascending_flags)
        : AbstractPlanNode(PlanType::SORT), order_(std::move(order_indices)), asc_(std::move(ascending_flags))
    {
    }
    std::unique_ptr<AbstractOperator> createOperator(ExecutionContext &context) const override
    {
        return std::make_unique<SortOperator>(context, output_schema_.clone(), child_op(context, 0, 1, "SortPlanNode"),
                                              order_, asc_);
    }
    std::string toString() const override
    {
        std::string s = "Sort(";
        for (size_t i = 0; i < order_.size(); ++i)
            s += (i ? ", #" : "#") + std::to_string(order_[i]) + (i < asc_.size() && !asc_[i] ? " DESC" : " ASC");
        return s + ")";
    }

private:
    std::vector<size_t> order_;
    std::vector<bool> asc_;
};

// FilterCompactionPlanNode (include/planner/filter_compaction_plan_node.hpp:14-27,
// src/planner/filter_compaction_plan_node.cpp): compacts the rows whose mask
// byte (column mask_index) is set; setCompactColumns chooses the columns.
// The node selects either by a mask column or by a predicate evaluated on
// the device.
class FilterCompactionPlanNode : public AbstractPlanNode {
public:
    explicit FilterCompactionPlanNode(size_t mask_index) : AbstractPlanNode(PlanType::COMPACTION), mask_(mask_index) {}
    // the reference's form (filter_compaction_plan_node.cpp:17-28): the mask
    // is the output schema's "$_mask" column
    FilterCompactionPlanNode(Schema output_schema, std::vector<bool> compact_columns = {})
        : AbstractPlanNode(PlanType::COMPACTION), mask_(kByName), compact_(std::move(compact_columns))
    {
        output_schema_ = std::move(output_schema);
    }
    FilterCompactionPlanNode(std::vector<PredicateTerm> predicate, ConnectiveType connective = ConnectiveType::AND)
        : AbstractPlanNode(PlanType::COMPACTION), mask_(kNoMask), pred_(std::move(predicate)), conn_(connective)
    {
    }
    void setCompactColumns(std::vector<bool> compact_columns) { compact_ = std::move(compact_columns); }
    const std::vector<bool> &getCompactColumns() const { return compact_; }
    std::unique_ptr<AbstractOperator> createOperator(ExecutionContext &context) const override
    {
        auto child = child_op(context, 0, 1, "FilterCompactionPlanNode");
        if (mask_ == kByName)
            return std::make_unique<FilterCompactionOperator>(context, output_schema_.clone(), std::move(child),
                                                              compact_);
        if (mask_ == kNoMask)
            return std::make_unique<FilterCompactionOperator>(context, std::move(child), pred_, conn_, compact_);
        return std::make_unique<FilterCompactionOperator>(context, std::move(child), mask_, compact_);
    }
    std::string toString() const override
    {
        if (mask_ == kByName) return "FilterCompaction()";
        if (mask_ != kNoMask) return "FilterCompaction(mask #" + std::to_string(mask_) + ")";
        static const char *const ops[] = {"=", "<>", "<", "<=", ">", ">="};
        std::string s = "FilterCompaction(";
        for (size_t i = 0; i < pred_.size(); ++i) {
            if (i) s += conn_ == ConnectiveType::AND ? " AND " : " OR ";
            s += "#" + std::to_string(pred_[i].column) + " " + ops[static_cast<int>(pred_[i].cmp)] + " c";
        }
        return s + ")";
    }

private:
    static constexpr size_t kNoMask = ~(size_t)0, kByName = ~(size_t)1;
    size_t mask_;
    std::vector<PredicateTerm> pred_;
    ConnectiveType conn_ = ConnectiveType::AND;
    std::vector<bool> compact_;
};

// SortMergeJoinPlanNode (include/planner/sort_merge_join_plan_node.hpp:13-35,
// src/planner/sort_merge_join_plan_node.cpp): children are (left, right).
class SortMergeJoinPlanNode : public AbstractPlanNode {
public:
    SortMergeJoinPlanNode(std::pair<size_t, size_t> join_key_indices, JoinType join_type = JoinType::INNER,
                          bool oblivious_materialization = false)
        : AbstractPlanNode(PlanType::SORT_MERGE_JOIN), keys_(join_key_indices), type_(join_type),
          oblivious_(oblivious_materialization)
    {
    }
    // the reference's form (sort_merge_join_plan_node.cpp:26-41)
    SortMergeJoinPlanNode(Schema output_schema, std::pair<size_t, size_t> join_key_indices,
                          std::vector<const Table *> left_source_tables, std::vector<const Table *> right_source_tables,
                          JoinType join_type = JoinType::INNER)
        : AbstractPlanNode(PlanType::SORT_MERGE_JOIN), keys_(join_key_indices), type_(join_type), oblivious_(false),
          left_tables_(std::move(left_source_tables)), right_tables_(std::move(right_source_tables))
    {
        output_schema_ = std::move(output_schema);
    }
    JoinType getJoinType() const { return type_; }
    std::unique_ptr<AbstractOperator> createOperator(ExecutionContext &context) const override
    {
        auto left = child_op(context, 0, 2, "SortMergeJoinPlanNode");
        auto right = child_op(context, 1, 2, "SortMergeJoinPlanNode");
        return std::make_unique<SortMergeJoinOperator>(context, output_schema_.clone(), std::move(left),
                                                       std::move(right), keys_, left_tables_, right_tables_, type_,
                                                       oblivious_);
    }
    std::string toString() const override
    {
        return "SortMergeJoin(#" + std::to_string(keys_.first) + " = #" + std::to_string(keys_.second) + ")";
    }

private:
    std::pair<size_t, size_t> keys_;
    JoinType type_;
    bool oblivious_;
    std::vector<const Table *> left_tables_, right_tables_;
};

// MaterializationPlanNode (include/planner/materialization_plan_node.hpp,
// src/planner/materialization_plan_node.cpp)
class MaterializationPlanNode : public AbstractPlanNode {
public:
    explicit MaterializationPlanNode(std::vector<MaterializedColumn> columns)
        : AbstractPlanNode(PlanType::MATERIALIZATION), cols_(std::move(columns))
    {
    }
    std::unique_ptr<AbstractOperator> createOperator(ExecutionContext &context) const override
    {
        return std::make_unique<MaterializationOperator>(context, child_op(context, 0, 1, "MaterializationPlanNode"),
                                                         cols_);
    }
    std::string toString() const override { return "Materialization(" + std::to_string(cols_.size()) + " columns)"; }

private:
    std::vector<MaterializedColumn> cols_;
};

// PK/FK join + SUM GROUP BY; children are (dimension, fact).
class PKFKJoinGroupSumPlanNode : public AbstractPlanNode {
public:
    PKFKJoinGroupSumPlanNode(size_t dim_key, size_t dim_group, size_t fact_fk, size_t fact_val, size_t max_groups,
                             bool dense_groups)
        : AbstractPlanNode(PlanType::AGGREGATE), dk_(dim_key), dg_(dim_group), fk_(fact_fk), fv_(fact_val),
          max_groups_(max_groups), dense_(dense_groups)
    {
    }
    std::unique_ptr<AbstractOperator> createOperator(ExecutionContext &context) const override
    {
        auto dim = child_op(context, 0, 2, "PKFKJoinGroupSumPlanNode");
        auto fact = child_op(context, 1, 2, "PKFKJoinGroupSumPlanNode");
        return std::make_unique<PKFKJoinGroupSumOperator>(context, std::move(dim), std::move(fact), dk_, dg_, fk_, fv_,
                                                          max_groups_, dense_);
    }
    std::string toString() const override
    {
        return "PKFKJoinGroupSum(#" + std::to_string(dk_) + " = #" + std::to_string(fk_) + ", group #" +
               std::to_string(dg_) + ", sum #" + std::to_string(fv_) + ")";
    }

private:
    size_t dk_, dg_, fk_, fv_, max_groups_;
    bool dense_;
};

// The planner's sort-merge join wiring (src/planner/query_planner.cpp:443-476):
// each input sorted ascending on its key, then the join of the two sorted
// inputs.
inline std::unique_ptr<AbstractPlanNode> planSortMergeJoin(std::unique_ptr<AbstractPlanNode> left,
                                                           std::unique_ptr<AbstractPlanNode> right,
                                                           std::pair<size_t, size_t> join_key_indices,
                                                           bool oblivious_materialization = false)
{
    auto ls = std::make_unique<SortPlanNode>(std::vector<size_t>{join_key_indices.first}, std::vector<bool>{true});
    ls->addChild(std::move(left));
    auto rs = std::make_unique<SortPlanNode>(std::vector<size_t>{join_key_indices.second}, std::vector<bool>{true});
    rs->addChild(std::move(right));
    auto j = std::make_unique<SortMergeJoinPlanNode>(join_key_indices, JoinType::INNER, oblivious_materialization);
    j->addChild(std::move(ls));
    j->addChild(std::move(rs));
    return j;
}

// The planner's single-table and ORDER BY wiring (query_planner.cpp:247-252,
// 518-544): compaction of the scan by its mask column, then, with order
// columns given, a sort over the compacted rows.
inline std::unique_ptr<AbstractPlanNode> planFilterOrderBy(std::unique_ptr<AbstractPlanNode> scan, size_t mask_index,
                                                           std::vector<size_t> order_indices = {},
                                                           std::vector<bool> ascending_flags = {})
{
    auto f = std::make_unique<FilterCompactionPlanNode>(mask_index);
    f->addChild(std::move(scan));
    if (order_indices.empty()) return f;
    auto s = std::make_unique<SortPlanNode>(std::move(order_indices), std::move(ascending_flags));
    s->addChild(std::move(f));
    return s;
}

// ------------------------------------------------- oblivious position maps

// A table's position map (record id -> position), re-shuffled after every
// query that read the table (ObliviousTable, src/data/oblivious_table.cpp).
// The reference draws sigma with a host Fisher-Yates over std::mt19937_64,
// uploads it and composes new[i] = old[sigma[i]] on the device, a gather
// whose addresses follow sigma (:175-205, oblivious_shuffle.cuh:10-21).  Here
// the map itself goes through OP_SHUFFLE: Philox4x32-10 tags of a fresh
// 64-bit seed sorted by the bitonic network with the map entries as
// payload, so new = old composed with a uniform permutation, computed without
// a host permutation, an upload or a data-dependent address.  The map has a
// power-of-two capacity; rows >= getRowCount() are not part of it.
class ObliviousTable {
public:
    ObliviousTable(std::string name, const RowBatch *source, TaskManager &tm,
                   uint64_t seed = std::random_device{}())
        : name_(std::move(name)), source_(source), tm_(tm), rows_(source ? source->getRowCount() : 0), rng_(seed)
    {
        if (rows_ >= (1ull << 32)) throw ExecutionError("ObliviousTable: more than 2^32 rows");
        pad_ = 1;
        while (pad_ < rows_) pad_ <<= 1;
        hip_check(hipMalloc(reinterpret_cast<void **>(&d_active_), pad_ * sizeof(uint32_t)), "hipMalloc");
        hip_check(hipMalloc(reinterpret_cast<void **>(&d_pending_), pad_ * sizeof(uint32_t)), "hipMalloc");
        // identity, then the initial shuffle (:22-33)
        h_map_.resize(pad_);
        for (size_t i = 0; i < pad_; ++i) h_map_[i] = static_cast<uint32_t>(i);
        hip_check(hipMemcpy(d_active_, h_map_.data(), pad_ * sizeof(uint32_t), hipMemcpyHostToDevice), "hipMemcpy");
        vo_shuffle_args a{d_active_, nullptr, pad_, rng_(), rows_};
        if (rows_) tm_.check(vo_shuffle(tm_.native(), &a));
        pull();
    }
    ObliviousTable(const ObliviousTable &) = delete;
    ObliviousTable &operator=(const ObliviousTable &) = delete;
    ~ObliviousTable()
    {
        try {
            waitForShuffle();
        } catch (...) {
        }
        (void)hipFree(d_active_);
        (void)hipFree(d_pending_);
    }

    const std::string &getName() const { return name_; }
    size_t getRowCount() const { return rows_; }
    const RowBatch *getSourceTable() const { return source_; }
    uint32_t getPosition(uint32_t record_id) const { return h_map_.at(record_id); }
    uint32_t *getPositionMapDevice() const { return d_active_; }
    // upload the host map (after host-side edits), as :152-165
    void syncPositionMapToGPU()
    {
        if (!shuffling_.load())
            hip_check(hipMemcpy(d_active_, h_map_.data(), pad_ * sizeof(uint32_t), hipMemcpyHostToDevice),
                      "hipMemcpy");
    }

    // pending = active, shuffled through the command queue; the active map
    // stays readable until waitForShuffle swaps them (:171-206)
    void startAsyncShuffle()
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (shuffling_.load()) return;
        hip_check(hipMemcpy(d_pending_, d_active_, pad_ * sizeof(uint32_t), hipMemcpyDeviceToDevice), "hipMemcpy");
        vo_command c{};
        c.opcode = VO_OP_SHUFFLE;
        c.args.shuffle = {d_pending_, nullptr, pad_, rng_(), rows_};
        pending_id_ = rows_ ? tm_.submitCommand(c) : 0;
        shuffling_.store(true);
        ++shuffles_;
    }
    void waitForShuffle()
    {
        if (!shuffling_.load()) return;
        std::lock_guard<std::mutex> lk(mu_);
        if (!shuffling_.load()) return;
        if (pending_id_) tm_.waitCommand(pending_id_);
        std::swap(d_active_, d_pending_);
        pull();
        shuffling_.store(false);
    }
    bool isShuffling() const { return shuffling_.load(); }
    size_t getShuffleCount() const { return shuffles_; }

private:
    void pull()
    {
        hip_check(hipMemcpy(h_map_.data(), d_active_, pad_ * sizeof(uint32_t), hipMemcpyDeviceToHost), "hipMemcpy");
    }

    std::string name_;
    const RowBatch *source_;
    TaskManager &tm_;
    size_t rows_, pad_ = 1;
    std::vector<uint32_t> h_map_;
    uint32_t *d_active_ = nullptr, *d_pending_ = nullptr;
    std::atomic<bool> shuffling_{false};
    uint64_t pending_id_ = 0;
    std::mutex mu_;
    std::mt19937_64 rng_; // seeds of the device shuffles
    size_t shuffles_ = 0;
};

// Query lifecycle over the registered tables (ObliviousTableManager,
// src/data/oblivious_table_manager.cpp): beginQuery waits for the pending
// shuffles of the tables a query names, endQuery starts a shuffle of every
// table marked accessed, enrichPositions maps record ids to positions with
// OP_ENRICH_POSITIONS.
class ObliviousTableManager {
public:
    struct Stats {
        size_t total_shuffles = 0;
        size_t tables_registered = 0;
    };

    explicit ObliviousTableManager(TaskManager &tm) : tm_(tm) {}
    ~ObliviousTableManager()
    {
        try {
            syncAllShuffles();
        } catch (...) {
        }
    }

    void registerTable(const std::string &name, const RowBatch *source, uint64_t seed = std::random_device{}())
    {
        std::lock_guard<std::mutex> lk(tables_mu_);
        tables_[name] = std::make_unique<ObliviousTable>(name, source, tm_, seed);
        std::lock_guard<std::mutex> sl(stats_mu_);
        stats_.tables_registered++;
    }
    bool hasTable(const std::string &name) const
    {
        std::lock_guard<std::mutex> lk(tables_mu_);
        return tables_.count(name) != 0;
    }
    ObliviousTable &getTable(const std::string &name)
    {
        std::lock_guard<std::mutex> lk(tables_mu_);
        auto it = tables_.find(name);
        if (it == tables_.end()) throw std::runtime_error("Oblivious table not found: " + name);
        return *it->second;
    }

    void beginQuery(const std::vector<std::string> &table_names)
    {
        std::lock_guard<std::mutex> lk(query_mu_);
        accessed_.clear();
        for (const auto &name : table_names)
            if (hasTable(name)) {
                ObliviousTable &t = getTable(name);
                if (t.isShuffling()) t.waitForShuffle();
                t.syncPositionMapToGPU();
            }
    }
    void markAccessed(const std::string &name)
    {
        std::lock_guard<std::mutex> lk(query_mu_);
        accessed_.insert(name);
    }
    void endQuery()
    {
        std::lock_guard<std::mutex> lk(query_mu_);
        for (const auto &name : accessed_)
            if (hasTable(name)) {
                getTable(name).startAsyncShuffle();
                std::lock_guard<std::mutex> sl(stats_mu_);
                stats_.total_shuffles++;
            }
    }
    // finish the background shuffles, then one timed shuffle of every table
    // the last query accessed (:157-181); the times are kept per table
    void profileAccessedTableShuffles()
    {
        std::lock_guard<std::mutex> lk(query_mu_);
        syncAllShuffles();
        for (const auto &name : accessed_) {
            if (!hasTable(name)) continue;
            ObliviousTable &t = getTable(name);
            const auto t0 = std::chrono::steady_clock::now();
            t.startAsyncShuffle();
            t.waitForShuffle();
            const auto us =
                std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count();
            std::lock_guard<std::mutex> sl(stats_mu_);
            shuffle_us_[name] = static_cast<double>(us);
        }
    }
    double controlledShuffleMicros(const std::string &name) const
    {
        std::lock_guard<std::mutex> sl(stats_mu_);
        auto it = shuffle_us_.find(name);
        return it == shuffle_us_.end() ? -1.0 : it->second;
    }

    void enrichPositions(const std::string &table_name, const uint32_t *d_record_ids, uint32_t *d_positions, size_t n)
    {
        ObliviousTable &t = getTable(table_name);
        vo_command c{};
        c.opcode = VO_OP_ENRICH_POSITIONS;
        c.args.enrich_positions = {d_record_ids, t.getPositionMapDevice(), d_positions, n};
        tm_.waitCommand(tm_.submitCommand(c));
    }

    void syncAllShuffles()
    {
        std::lock_guard<std::mutex> lk(tables_mu_);
        for (auto &kv : tables_)
            if (kv.second->isShuffling()) kv.second->waitForShuffle();
    }
    Stats getStats() const
    {
        std::lock_guard<std::mutex> sl(stats_mu_);
        return stats_;
    }

private:
    TaskManager &tm_;
    std::map<std::string, std::unique_ptr<ObliviousTable>> tables_;
    mutable std::mutex tables_mu_;
    std::set<std::string> accessed_;
    std::mutex query_mu_;
    mutable std::mutex stats_mu_;
    Stats stats_;
    std::map<std::string, double> shuffle_us_;
};

} // namespace velodb_amd
