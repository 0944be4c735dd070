// spira_amd.hpp -- C++23 host side of the device backend.
//
// spira::matrix / spira::parallel_matrix for <soa_tag, uint32_t, double>
// and their spmv overloads, with lock() and spmv running on an MI355X through
// the C-ABI of spira_amd.h (libspira_amd.so).  Names, argument meaning, check
// order and exception types follow the reference:
//   matrix           include/spira/matrix/matrix.hpp:30-178 (interface),
//                    :277-357 (lock/open), :367-443 (queries), :453-491
//                    (insert/clear/row_at_mut), :520-605 (load_csr)
//   parallel_matrix  include/spira/parallel/parallel_matrix.hpp (interface,
//                    uniform split :296-306, owner() :236-242, lock
//                    :393-405, rebalance :434-497, partition_at :210-222)
//   spmv             include/spira/serial/spmv.hpp:142-186,
//                    include/spira/parallel/algorithms/spmv.hpp:127-161
// Open-mode inserts stay on the host in insertion order (the row buffers of
// row.hpp:162-169); lock() hands them to the device, which deduplicates them
// (last write wins, zeros kept as deletions) and builds or merges the CSR.
// The committed CSR lives in HBM; csr() downloads a host copy on demand.
//
// Header-only; link with -lspira_amd.  Needs C++23 (-std=c++2b).
#ifndef SPIRA_AMD_HPP
#define SPIRA_AMD_HPP

#include "spira_amd.h"

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <exception>
#include <memory>
#include <new>
#include <optional>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

namespace spira_amd {

// config::matrix_mode (config.hpp:14)
enum class matrix_mode : uint8_t { open, locked };

// Status code -> the reference's exception classes.
inline void throw_status(int rc, const char *msg)
{
    const std::string m = msg ? msg : "";
    switch (rc) {
    case SPIRA_OK: return;
    case SPIRA_ERR_INVALID_ARGUMENT: throw std::invalid_argument(m);
    case SPIRA_ERR_LOGIC: throw std::logic_error(m);
    case SPIRA_ERR_OUT_OF_RANGE: throw std::out_of_range(m);
    case SPIRA_ERR_NOMEM: throw std::bad_alloc();
    case SPIRA_ERR_UNSUPPORTED: throw std::length_error(m);
    default: throw std::runtime_error(m.empty() ? "spira_amd: device error" : m);
    }
}

// One GPU and its stream (spira_dev).
class device {
public:
    explicit device(int ordinal = 0, void *stream = nullptr)
    {
        spira_dev *d = nullptr;
        const int rc = spira_dev_create(&d, ordinal, stream);
        if (rc != SPIRA_OK) throw_status(rc, spira_dev_last_error(nullptr));
        dev_ = d;
    }
    ~device() { spira_dev_destroy(dev_); }
    device(const device &) = delete;
    device &operator=(const device &) = delete;
    [[nodiscard]] spira_dev *get() const noexcept { return dev_; }
    void check(int rc) const
    {
        if (rc != SPIRA_OK) throw_status(rc, spira_dev_last_error(dev_));
    }
    void synchronize() const { check(spira_dev_synchronize(dev_)); }

private:
    spira_dev *dev_ = nullptr;
};

// Process-wide device 0, created on first use.
inline std::shared_ptr<device> default_device()
{
    static std::shared_ptr<device> d = std::make_shared<device>(0);
    return d;
}

// The observable surface of csr_storage<soa_tag, uint32_t, double>
// (csr_storage_soa.hpp:17-76): offsets[n_rows+1], cols[nnz], vals[nnz].
struct csr_storage {
    std::size_t n_rows{0};
    std::size_t nnz{0};
    std::vector<std::size_t> offsets;
    std::vector<uint32_t> cols;
    std::vector<double> vals;
    [[nodiscard]] bool is_built() const noexcept { return !offsets.empty(); }
};

// A locked device CSR (spira_dcsr), owned.
class device_csr {
public:
    device_csr(std::shared_ptr<device> dev, std::size_t n_rows, std::size_t n_cols) : dev_(std::move(dev))
    {
        spira_dcsr *a = nullptr;
        dev_->check(spira_dcsr_create(dev_->get(), n_rows, n_cols, &a));
        a_ = a;
    }
    ~device_csr() { spira_dcsr_destroy(a_); }
    device_csr(const device_csr &) = delete;
    device_csr &operator=(const device_csr &) = delete;
    [[nodiscard]] spira_dcsr *get() const noexcept { return a_; }
    [[nodiscard]] const device &dev() const noexcept { return *dev_; }
    [[nodiscard]] std::size_t nnz() const
    {
        uint64_t z = 0;
        dev_->check(spira_dcsr_info(a_, nullptr, nullptr, &z, nullptr));
        return (std::size_t)z;
    }
    [[nodiscard]] bool built() const
    {
        int b = 0;
        dev_->check(spira_dcsr_info(a_, nullptr, nullptr, nullptr, &b));
        return b != 0;
    }

private:
    std::shared_ptr<device> dev_;
    spira_dcsr *a_ = nullptr;
};

class matrix {
public:
    using index_type = uint32_t;
    using value_type = double;
    using size_type = std::size_t;
    using shape_type = std::pair<size_type, size_type>;

    explicit matrix(size_type row_limit, size_type column_limit, std::shared_ptr<device> dev = default_device())
        : rows_(row_limit), cols_(column_limit), dev_(std::move(dev)), buf_(row_limit), dirty_(row_limit, 0),
          csr_(std::make_unique<device_csr>(dev_, row_limit, column_limit))
    {
    }
    virtual ~matrix() = default;
    matrix(matrix &&) noexcept = default;
    matrix &operator=(matrix &&) noexcept = default;

    // shape
    [[nodiscard]] shape_type shape() const noexcept { return {rows_, cols_}; }
    [[nodiscard]] size_type n_rows() const noexcept { return rows_; }
    [[nodiscard]] size_type n_cols() const noexcept { return cols_; }

    // modes
    [[nodiscard]] matrix_mode mode() const noexcept { return mode_; }
    [[nodiscard]] bool is_locked() const noexcept { return mode_ == matrix_mode::locked; }
    [[nodiscard]] bool is_open() const noexcept { return mode_ == matrix_mode::open; }

    // matrix.hpp:277-342: no-op when locked; dedup + build/merge on the device.
    void lock() { lock_with(dirty_.data()); }

    // matrix.hpp:348-357: the committed CSR is kept as the next merge base.
    void open() noexcept { mode_ = matrix_mode::open; }

    // queries
    [[nodiscard]] size_type row_nnz(index_type r) const
    {
        validate_row_index(r);
        const auto *h = csr();
        const size_type committed = h ? h->offsets[r + 1] - h->offsets[r] : 0;
        return committed + (is_open() ? buf_[r].size() : 0);
    }
    [[nodiscard]] bool empty() const noexcept { return nnz() == 0; }
    [[nodiscard]] size_type nnz() const noexcept
    {
        size_type total = csr_built_ ? committed_nnz_ : 0;
        if (is_open())
            for (const auto &b : buf_) total += b.size();
        return total;
    }
    // matrix.hpp:88-91: nullptr before the first lock.
    [[nodiscard]] const csr_storage *csr() const
    {
        if (!csr_built_) return nullptr;
        if (!host_) {
            csr_storage h;
            h.n_rows = rows_;
            h.nnz = csr_->nnz();
            h.offsets.resize(rows_ + 1);
            h.cols.resize(std::max<size_type>(h.nnz, 1));
            h.vals.resize(std::max<size_type>(h.nnz, 1));
            static_assert(sizeof(std::size_t) == sizeof(uint64_t));
            check(spira_dcsr_download_f64(csr_->get(), reinterpret_cast<uint64_t *>(h.offsets.data()),
                                          h.cols.data(), h.vals.data(), 0));
            h.cols.resize(h.nnz);
            h.vals.resize(h.nnz);
            host_ = std::move(h);
        }
        return &*host_;
    }
    [[nodiscard]] bool contains(index_type r, index_type c) const
    {
        validate_row_index(r);
        validate_col_index(c);
        if (is_open() && buffered(r, c)) return true;
        return committed(r, c).has_value();
    }
    [[nodiscard]] value_type get(index_type r, index_type c) const
    {
        validate_row_index(r);
        validate_col_index(c);
        if (is_open())
            if (auto v = buffered(r, c)) return *v;  // buffer first: the last write wins
        return committed(r, c).value_or(0.0);
    }
    // Sum of the row's values, with open-mode inserts applied (last write wins).
    [[nodiscard]] value_type accumulate(index_type r) const
    {
        validate_row_index(r);
        std::vector<std::pair<index_type, value_type>> row;
        if (const auto *h = csr())
            for (size_type k = h->offsets[r]; k < h->offsets[r + 1]; ++k) row.emplace_back(h->cols[k], h->vals[k]);
        if (is_open()) {
            for (const auto &[c, v] : buf_[r]) {
                auto it = std::find_if(row.begin(), row.end(), [c](const auto &e) { return e.first == c; });
                if (it != row.end()) it->second = v;
                else row.emplace_back(c, v);
            }
        }
        value_type s = 0.0;
        for (const auto &e : row) s += e.second;
        return s;
    }

    // mutation (open mode)
    void insert(index_type r, index_type c, const value_type &v)
    {
        if (mode_ != matrix_mode::open) throw std::logic_error(insert_error());
        validate_row_index(r);
        validate_col_index(c);
        buf_[r].emplace_back(c, v);
        dirty_[r] = 1;
    }
    void clear()
    {
        if (mode_ != matrix_mode::open) throw std::logic_error(clear_error());
        for (auto &b : buf_) b.clear();
        std::fill(dirty_.begin(), dirty_.end(), uint8_t{0});
    }
    // The effect of row_at_mut() (matrix.hpp:483-491) on the next lock: the
    // row counts as dirty even without an insert.
    void mark_dirty(index_type r)
    {
        if (mode_ != matrix_mode::open) throw std::logic_error("matrix::row_at_mut() requires open mode");
        validate_row_index(r);
        dirty_[r] = 1;
    }

    // matrix.hpp:520-605: validated on the device; the matrix ends up locked.
    void load_csr(csr_storage &&h)
    {
        if (h.offsets.size() != rows_ + 1)
            throw std::invalid_argument("spira::matrix::load_csr: CSR row count does not match matrix dimensions");
        if (h.cols.size() != h.vals.size()) throw std::invalid_argument("load_csr: cols and vals differ in length");
        check(spira_dcsr_upload_f64(csr_->get(), h.cols.size(), reinterpret_cast<const uint64_t *>(h.offsets.data()),
                                    h.cols.data(), h.vals.data(), 0));
        for (auto &b : buf_) b.clear();
        std::fill(dirty_.begin(), dirty_.end(), uint8_t{0});
        mode_ = matrix_mode::locked;
        csr_built_ = true;
        committed_nnz_ = h.cols.size();
        host_.reset();
    }

    // y = A * x (matrix_operators.hpp:88-96 / operator* of matrix.hpp:146).
    std::vector<value_type> operator*(const std::vector<value_type> &x) const
    {
        std::vector<value_type> y(rows_);
        spmv_into(x, y, "The size of the input vector x does not match the number of columns of the matrix",
                  "The size of the output vector y does not match the number of rows of the matrix");
        return y;
    }

    // the device objects
    [[nodiscard]] spira_dcsr *device_handle() const noexcept { return csr_->get(); }
    [[nodiscard]] const device &dev() const noexcept { return *dev_; }
    [[nodiscard]] std::shared_ptr<device> device_ptr() const noexcept { return dev_; }

    // A^T into out (a matrix of the swapped shape, or *this when square) on
    // the device; out ends locked with the transposed CSR committed and no
    // pending inserts.  flags: SPIRA_TRANSPOSE_DROP_ZEROS for the parallel form.
    void transpose_into(matrix &out, uint32_t flags) const
    {
        check(spira_transpose_f64(csr_->get(), out.csr_->get(), flags));
        out.adopt_device_result();
    }

    // out = *this + B (flags: SPIRA_ADD_SUBTRACT for *this - B) on the device;
    // out (this shape; may be *this or B) ends locked with no pending inserts.
    void add_into(const matrix &B, matrix &out, uint32_t flags) const
    {
        check(spira_add_f64(csr_->get(), B.csr_->get(), out.csr_->get(), flags));
        out.adopt_device_result();
    }

    // out = *this * B on the device; out has this matrix's rows and B's columns.
    void spgemm_into(const matrix &B, matrix &out) const
    {
        check(spira_spgemm_f64(csr_->get(), B.csr_->get(), out.csr_->get()));
        out.adopt_device_result();
    }

    // operator+ / operator+= (matrix_operators.hpp:21-47): MatrixAddition.
    matrix operator+(const matrix &other) const { return add_checked(other, 0, "Matrices aren't the same size"); }
    matrix &operator+=(const matrix &other)
    {
        *this = add_checked(other, 0, "Matrices aren't the same size");
        return *this;
    }
    // operator- / operator-= (matrix_operators.hpp:29-62): the shape check,
    // then MatrixAddition(*this, other * -1).
    matrix operator-(const matrix &other) const
    {
        return add_checked(other, SPIRA_ADD_SUBTRACT, "operator-: matrix shapes must match");
    }
    matrix &operator-=(const matrix &other)
    {
        *this = add_checked(other, SPIRA_ADD_SUBTRACT, "operator-=: matrix shapes must match");
        return *this;
    }
    // operator* / operator*= of two matrices (matrix_operators.hpp:67-83): spgemm.
    matrix operator*(const matrix &other) const { return spgemm_checked(other, "A.cols must equal B.rows"); }
    matrix &operator*=(const matrix &other)
    {
        *this = spgemm_checked(other, "A.cols must equal B.rows");
        return *this;
    }

    // the checks of serial::algorithms::MatrixAddition (serial/
    // matrix_addition.hpp:88-101; A - B checks that B is locked in its scaled
    // copy first, scalars.hpp:26-27) and of spgemm (serial/spgemm.hpp:14-23).
    matrix add_checked(const matrix &other, uint32_t flags, const char *shape_msg) const
    {
        if (shape() != other.shape()) throw std::invalid_argument(shape_msg);
        if ((flags & SPIRA_ADD_SUBTRACT) && !other.is_locked())
            throw std::logic_error("multiplication_scaler: input matrix must be locked");
        if (!is_locked()) throw std::logic_error("MatrixAddition: A must be locked");
        if (!other.is_locked()) throw std::logic_error("MatrixAddition: B must be locked");
        matrix out(rows_, cols_, dev_);
        add_into(other, out, flags);
        return out;
    }
    matrix spgemm_checked(const matrix &other, const char *shape_msg) const
    {
        if (cols_ != other.rows_) throw std::invalid_argument(shape_msg);
        if (!is_locked()) throw std::logic_error("spgemm: A must be locked");
        if (!other.is_locked()) throw std::logic_error("spgemm: B must be locked");
        matrix out(rows_, other.cols_, dev_);
        spgemm_into(other, out);
        return out;
    }

    // checks in serial/spmv.hpp:150-158 order, then the device SpMV
    void spmv_into(const std::vector<value_type> &x, std::vector<value_type> &y, const char *x_msg,
                   const char *y_msg, int variant = SPIRA_SPMV_AUTO) const
    {
        if (x.size() != cols_) throw std::invalid_argument(x_msg);
        if (y.size() != rows_) throw std::invalid_argument(y_msg);
        if (!is_locked()) throw std::logic_error("spmv: matrix must be locked");
        if (!csr_built_) {  // locked without a CSR: every row is empty
            std::fill(y.begin(), y.end(), 0.0);
            return;
        }
        check(spira_spmv_f64(csr_->get(), x.data(), x.size(), y.data(), y.size(), variant,
                             reproducible_ ? SPIRA_REPRODUCIBLE : SPIRA_PTR_HOST));
    }

    // The float instantiation's SpMV (FP32 x and y; the values, inserted as
    // floats, are exact in the device's FP64 storage), same checks.
    void spmv_into(const std::vector<float> &x, std::vector<float> &y, const char *x_msg, const char *y_msg,
                   int variant = SPIRA_SPMV_AUTO) const
    {
        if (x.size() != cols_) throw std::invalid_argument(x_msg);
        if (y.size() != rows_) throw std::invalid_argument(y_msg);
        if (!is_locked()) throw std::logic_error("spmv: matrix must be locked");
        if (!csr_built_) {
            std::fill(y.begin(), y.end(), 0.0f);
            return;
        }
        check(spira_spmv_f32(csr_->get(), x.data(), x.size(), y.data(), y.size(), variant, SPIRA_PTR_HOST));
    }

    // Bit-identical y on every SpMV with the same matrix and x (the kernel
    // choice keeps to variants whose summation order is fixed), as the
    // reference's per-row loop gives; off by default (the fastest variant).
    void set_reproducible(bool on) { reproducible_ = on; }
    bool reproducible() const { return reproducible_; }

protected:
    void check(int rc) const { dev_->check(rc); }
    virtual const char *insert_error() const { return "matrix::insert() requires open mode"; }
    virtual const char *clear_error() const { return "matrix::clear() requires open mode"; }
    virtual void validate_row_index(index_type r) const
    {
        if ((size_type)r >= rows_) throw std::out_of_range("spira::matrix: row_index out of range");
    }
    virtual void validate_col_index(index_type c) const
    {
        if ((size_type)c >= cols_) throw std::out_of_range("spira::matrix: col_index out of range");
    }

    // Stage the row buffers (insertion order kept) and lock on the device.
    // dirty == nullptr is the parallel rule: a row is dirty iff it has inserts.
    // The device wrote a new committed CSR into this matrix's handle.
    void adopt_device_result()
    {
        for (auto &b : buf_) b.clear();
        std::fill(dirty_.begin(), dirty_.end(), uint8_t{0});
        mode_ = matrix_mode::locked;
        csr_built_ = true;
        committed_nnz_ = csr_->nnz();
        host_.reset();
    }

    void lock_with(const uint8_t *dirty)
    {
        if (mode_ == matrix_mode::locked) return;
        std::vector<uint64_t> ptr(rows_ + 1, 0);
        for (size_type r = 0; r < rows_; ++r) ptr[r + 1] = ptr[r] + buf_[r].size();
        std::vector<uint32_t> col(ptr[rows_]);
        std::vector<double> val(ptr[rows_]);
        for (size_type r = 0; r < rows_; ++r) {
            uint64_t k = ptr[r];
            for (const auto &[c, v] : buf_[r]) {
                col[k] = c;
                val[k] = v;
                ++k;
            }
        }
        uint64_t nnz = 0;
        check(spira_lock_f64(csr_->get(), ptr.data(), col.data(), val.data(), csr_built_ ? dirty : nullptr,
                             SPIRA_PTR_HOST, &nnz));
        for (auto &b : buf_) b.clear();
        std::fill(dirty_.begin(), dirty_.end(), uint8_t{0});
        mode_ = matrix_mode::locked;
        csr_built_ = true;
        committed_nnz_ = nnz;
        host_.reset();
    }

    std::optional<value_type> buffered(index_type r, index_type c) const
    {
        const auto &b = buf_[r];
        for (auto it = b.rbegin(); it != b.rend(); ++it)
            if (it->first == c) return it->second;
        return std::nullopt;
    }
    std::optional<value_type> committed(index_type r, index_type c) const
    {
        const auto *h = csr();
        if (!h) return std::nullopt;
        const auto first = h->cols.begin() + (std::ptrdiff_t)h->offsets[r];
        const auto last = h->cols.begin() + (std::ptrdiff_t)h->offsets[r + 1];
        const auto it = std::lower_bound(first, last, c);
        if (it == last || *it != c) return std::nullopt;
        return h->vals[(size_type)(it - h->cols.begin())];
    }

    matrix_mode mode_{matrix_mode::open};
    size_type rows_, cols_;
    std::shared_ptr<device> dev_;
    std::vector<std::vector<std::pair<index_type, value_type>>> buf_;  // open-mode inserts per row
    std::vector<uint8_t> dirty_;                                       // matrix::dirty_
    std::unique_ptr<device_csr> csr_;
    bool csr_built_ = false;
    bool reproducible_ = false;
    size_type committed_nnz_ = 0;
    mutable std::optional<csr_storage> host_;  // downloaded copy of the committed CSR
};

// A partition's view (partition.hpp:39-57): its rows and its slice of the
// committed CSR, offsets rebased to zero.
struct partition {
    std::size_t row_start{0};
    std::size_t row_end{0};
    csr_storage csr{};
    [[nodiscard]] std::size_t size() const noexcept { return row_end - row_start; }
    [[nodiscard]] std::size_t local_row(std::size_t global_row) const noexcept { return global_row - row_start; }
};

// parallel_matrix: rows split into n_threads contiguous partitions; the
// device holds one concatenated CSR whose row ranges are the partitions.
class parallel_matrix : public matrix {
public:
    parallel_matrix(size_type n_rows, size_type n_cols, size_type n_threads,
                    std::shared_ptr<device> dev = default_device())
        : matrix(n_rows, n_cols, std::move(dev))
    {
        if (n_threads < 1) throw std::invalid_argument("parallel_matrix requires at least one thread");
        bounds_.resize(n_threads + 1);
        for (size_type t = 0; t <= n_threads; ++t) bounds_[t] = t * n_rows / n_threads;  // :296-306
    }
    [[nodiscard]] size_type n_threads() const noexcept { return bounds_.size() - 1; }

    // dirty = row buffer non-empty (parallel_matrix.hpp:351-355)
    void lock() { lock_with(nullptr); }

    // parallel_matrix.hpp:434-497: nnz-balanced boundaries, computed on the
    // device from the committed offsets; no entry moves.
    void rebalance()
    {
        if (!is_locked()) throw std::logic_error("rebalance: matrix must be locked");
        if (!csr_built_) return;
        std::vector<uint64_t> b(n_threads() + 1);
        check(spira_partition_boundaries(csr_->get(), n_threads(), b.data()));
        for (size_type t = 0; t < b.size(); ++t) bounds_[t] = (size_type)b[t];
    }

    // owner() (parallel_matrix.hpp:236-242): the partition holding a row
    [[nodiscard]] size_type owner(size_type row) const
    {
        return (size_type)(std::upper_bound(bounds_.begin() + 1, bounds_.end() - 1, row) - (bounds_.begin() + 1));
    }

    [[nodiscard]] partition partition_at(size_type t) const
    {
        if (t >= n_threads())
            throw std::out_of_range("parallel_matrix::partition_at: thread index out of range");
        partition p;
        p.row_start = bounds_[t];
        p.row_end = bounds_[t + 1];
        if (const auto *h = csr()) {
            const size_type a = p.row_start, z = p.row_end, base = h->offsets[a];
            p.csr.n_rows = z - a;
            p.csr.nnz = h->offsets[z] - base;
            p.csr.offsets.resize(z - a + 1);
            for (size_type i = a; i <= z; ++i) p.csr.offsets[i - a] = h->offsets[i] - base;
            p.csr.cols.assign(h->cols.begin() + (std::ptrdiff_t)base,
                              h->cols.begin() + (std::ptrdiff_t)h->offsets[z]);
            p.csr.vals.assign(h->vals.begin() + (std::ptrdiff_t)base,
                              h->vals.begin() + (std::ptrdiff_t)h->offsets[z]);
        }
        return p;
    }
    [[nodiscard]] const std::vector<size_type> &partition_bounds() const noexcept { return bounds_; }

    // One staged row as parallel_fill hands it out (row.hpp:162-168): insert()
    // checks the column and appends to the row's open-mode buffer.
    class row_ref {
    public:
        void insert(index_type col, const value_type &val)
        {
            if ((size_type)col >= m_->cols_) throw std::out_of_range("Column index out of range");
            m_->buf_[r_].emplace_back(col, val);
            m_->dirty_[r_] = 1;
        }
        [[nodiscard]] size_type size() const noexcept { return m_->buf_[r_].size(); }

    private:
        friend class parallel_matrix;
        row_ref(parallel_matrix *m, size_type r) : m_(m), r_(r) {}
        parallel_matrix *m_;
        size_type r_;
    };
    // A partition's rows: rows[i] is global row row_start + i.
    class partition_rows {
    public:
        [[nodiscard]] row_ref operator[](size_type i) const { return row_ref(m_, a_ + i); }
        [[nodiscard]] size_type size() const noexcept { return z_ - a_; }

    private:
        friend class parallel_matrix;
        partition_rows(parallel_matrix *m, size_type a, size_type z) : m_(m), a_(a), z_(z) {}
        parallel_matrix *m_;
        size_type a_, z_;
    };

    // parallel_matrix.hpp:166-189: f(rows, row_start, row_end, t) runs once per
    // partition, each on its own host thread, and stages inserts into that
    // partition's rows only (distinct rows: no locking).  The matrix stays
    // open; lock() then builds or merges on the device as after insert().
    // The first exception thrown by f is rethrown after every thread joined;
    // inserts staged before it stay staged.
    template <class Func>
    void parallel_fill(Func &&f)
    {
        if (!is_open()) throw std::logic_error("parallel_fill: matrix must be open");
        const size_type nt = n_threads();
        std::vector<std::exception_ptr> err(nt);
        std::vector<std::thread> th;
        th.reserve(nt);
        for (size_type t = 0; t < nt; ++t)
            th.emplace_back([this, &f, &err, t] {
                try {
                    partition_rows rows(this, bounds_[t], bounds_[t + 1]);
                    f(rows, bounds_[t], bounds_[t + 1], t);
                } catch (...) {
                    err[t] = std::current_exception();
                }
            });
        for (auto &x : th) x.join();
        for (auto &e : err)
            if (e) std::rethrow_exception(e);
    }

protected:
    const char *insert_error() const override { return "parallel_matrix::insert() requires open mode"; }
    const char *clear_error() const override { return "parallel_matrix::clear() requires open mode"; }
    void validate_row_index(index_type r) const override
    {
        if ((size_type)r >= rows_) throw std::out_of_range("parallel_matrix: row index out of range");
    }
    void validate_col_index(index_type c) const override
    {
        if ((size_type)c >= cols_) throw std::out_of_range("parallel_matrix: col index out of range");
    }

private:
    std::vector<size_type> bounds_;
};

namespace serial::algorithms {
// serial/spmv.hpp:142-186
inline void spmv(const matrix &mat, const std::vector<double> &x, std::vector<double> &y)
{
    mat.spmv_into(x, y, "The size of the input vector x does not match the number of columns of the matrix",
                  "The size of the output vector y does not match the number of rows of the matrix");
}

// serial/spmv.hpp:90-135 (the float instantiation)
inline void spmv(const matrix &mat, const std::vector<float> &x, std::vector<float> &y)
{
    mat.spmv_into(x, y, "The size of the input vector x does not match the number of columns of the matrix",
                  "The size of the output vector y does not match the number of rows of the matrix");
}

// serial/transpose.hpp:35-157: a new locked matrix holding A^T (built on the
// device; rows sorted, values as stored).
inline matrix transpose(const matrix &mat)
{
    if (!mat.is_locked()) throw std::logic_error("transpose: matrix must be locked");
    matrix out(mat.n_cols(), mat.n_rows(), mat.device_ptr());
    mat.transpose_into(out, 0);
    return out;
}

// serial/transpose.hpp:158-173: open, square; lock, transpose, open again.
inline void transpose_itself(matrix &mat);

// serial/matrix_addition.hpp:88-112
inline matrix MatrixAddition(const matrix &A, const matrix &B)
{
    return A.add_checked(B, 0, "Matrices aren't the same size");
}

// serial/spgemm.hpp:12-50
inline matrix spgemm(const matrix &A, const matrix &B) { return A.spgemm_checked(B, "A.cols must equal B.rows"); }

inline void transpose_itself(matrix &mat)
{
    if (!mat.is_open()) throw std::logic_error("transpose_itself: matrix must be open");
    const auto [r, c] = mat.shape();
    if (r != c) throw std::logic_error("in-place transpose requires square matrix");
    mat.lock();
    mat.transpose_into(mat, 0);
    mat.open();
}
}  // namespace serial::algorithms

namespace parallel::algorithms {
// parallel/algorithms/spmv.hpp:127-161 (one launch: the partitions are row
// ranges of the device CSR)
inline void spmv(const parallel_matrix &mat, const std::vector<double> &x, std::vector<double> &y)
{
    mat.spmv_into(x, y, "spmv: x size does not match matrix column count",
                  "spmv: y size does not match matrix row count");
}

// parallel/algorithms/spmv.hpp:85-119 (the float instantiation)
inline void spmv(const parallel_matrix &mat, const std::vector<float> &x, std::vector<float> &y)
{
    mat.spmv_into(x, y, "spmv: x size does not match matrix column count",
                  "spmv: y size does not match matrix row count");
}

// parallel/algorithms/transpose.hpp:37-69: the output (same thread count,
// uniform partitions) ends locked; its lock() would drop explicit zeros, and
// so does the device transpose here.
inline parallel_matrix transpose(parallel_matrix &mat)
{
    if (!mat.is_locked()) throw std::logic_error("transpose: matrix must be locked");
    parallel_matrix out(mat.n_cols(), mat.n_rows(), mat.n_threads(), mat.device_ptr());
    mat.transpose_into(out, SPIRA_TRANSPOSE_DROP_ZEROS);
    return out;
}

// parallel/algorithms/matrix_addition.hpp:24-99: same shape and thread count;
// C has A's thread count (uniform partitions) and ends locked.
inline parallel_matrix MatrixAddition(parallel_matrix &A, parallel_matrix &B)
{
    if (A.shape() != B.shape()) throw std::invalid_argument("MatrixAddition: matrices must have the same shape");
    if (A.n_threads() != B.n_threads())
        throw std::invalid_argument("MatrixAddition: matrices must have the same thread count");
    if (!A.is_locked()) throw std::logic_error("MatrixAddition: A must be locked");
    if (!B.is_locked()) throw std::logic_error("MatrixAddition: B must be locked");
    parallel_matrix C(A.n_rows(), A.n_cols(), A.n_threads(), A.device_ptr());
    A.add_into(B, C, 0);
    return C;
}

// parallel/algorithms/spgemm.hpp:24-69: C is A.n_rows() x B.n_cols() with A's
// thread count.
inline parallel_matrix spgemm(parallel_matrix &A, parallel_matrix &B)
{
    if (A.n_cols() != B.n_rows()) throw std::invalid_argument("spgemm: A.n_cols() must equal B.n_rows()");
    if (!A.is_locked()) throw std::logic_error("spgemm: A must be locked");
    if (!B.is_locked()) throw std::logic_error("spgemm: B must be locked");
    parallel_matrix C(A.n_rows(), B.n_cols(), A.n_threads(), A.device_ptr());
    A.spgemm_into(B, C);
    return C;
}
}  // namespace parallel::algorithms

}  // namespace spira_amd

#endif  // SPIRA_AMD_HPP
