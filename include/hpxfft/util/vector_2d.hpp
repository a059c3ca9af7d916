// vector_2d.hpp -- row-major 2-D container with the API of HPX-FFT's
// hpxfft::util::vector_2d (core/include/hpxfft/util/vector_2d.hpp:10-86),
// written as a plain RAII type with no HPX dependency.
//
// Same members and semantics: zero-filling (n_row, n_col) constructor,
// fill constructor, row(i) = data() + i*n_col, operator()(i,j), at(i,j)
// throwing std::runtime_error("out of range exception"), bitwise operator==,
// iterators, size/n_row/n_col, swap.  Differences (deliberate): the buffer is
// owned by a std::unique_ptr and freed, copy assignment copies (the reference's
// swaps with its source), and a moved-from object is empty.
#ifndef HPXFFT_GPU_VECTOR_2D_HPP
#define HPXFFT_GPU_VECTOR_2D_HPP

#include <algorithm>
#include <cstddef>
#include <memory>
#include <stdexcept>
#include <utility>

namespace hpxfft::util
{
template <typename T>
class vector_2d
{
  public:
    using iterator = T *;
    using const_iterator = const T *;

    vector_2d() = default;

    vector_2d(std::size_t n_row, std::size_t n_col) :
        n_row_(n_row),
        n_col_(n_col),
        size_(n_row * n_col),
        values_(size_ ? new T[size_]() : nullptr)
    {
    }

    vector_2d(std::size_t n_row, std::size_t n_col, const T &v) :
        n_row_(n_row),
        n_col_(n_col),
        size_(n_row * n_col),
        values_(size_ ? new T[size_] : nullptr)
    {
        std::fill(begin(), end(), v);
    }

    vector_2d(const vector_2d &src) :
        n_row_(src.n_row_),
        n_col_(src.n_col_),
        size_(src.size_),
        values_(src.size_ ? new T[src.size_] : nullptr)
    {
        std::copy(src.begin(), src.end(), begin());
    }

    vector_2d(vector_2d &&mv) noexcept { swap(*this, mv); }

    ~vector_2d() = default;

    vector_2d &operator=(const vector_2d &src)
    {
        if (this != &src)
        {
            vector_2d tmp(src);
            swap(*this, tmp);
        }
        return *this;
    }

    vector_2d &operator=(vector_2d &&mv) noexcept
    {
        vector_2d tmp(std::move(mv));
        swap(*this, tmp);
        return *this;
    }

    T &operator()(std::size_t i, std::size_t j) { return values_[i * n_col_ + j]; }
    const T &operator()(std::size_t i, std::size_t j) const { return values_[i * n_col_ + j]; }

    T &at(std::size_t i, std::size_t j)
    {
        if (i * n_col_ + j >= size_) throw std::runtime_error("out of range exception");
        return values_[i * n_col_ + j];
    }
    const T &at(std::size_t i, std::size_t j) const
    {
        if (i * n_col_ + j >= size_) throw std::runtime_error("out of range exception");
        return values_[i * n_col_ + j];
    }

    T *data() noexcept { return values_.get(); }
    const T *data() const noexcept { return values_.get(); }

    iterator begin() noexcept { return values_.get(); }
    const_iterator begin() const noexcept { return values_.get(); }
    iterator end() noexcept { return values_.get() + size_; }
    const_iterator end() const noexcept { return values_.get() + size_; }
    const_iterator cbegin() const noexcept { return begin(); }
    const_iterator cend() const noexcept { return end(); }

    iterator row(std::size_t i) noexcept { return values_.get() + i * n_col_; }
    const_iterator row(std::size_t i) const noexcept { return values_.get() + i * n_col_; }

    std::size_t size() const noexcept { return size_; }
    std::size_t n_row() const noexcept { return n_row_; }
    std::size_t n_col() const noexcept { return n_col_; }

    friend bool operator==(const vector_2d &lhs, const vector_2d &rhs)
    {
        if (lhs.n_row_ != rhs.n_row_ || lhs.n_col_ != rhs.n_col_) return false;
        for (std::size_t i = 0; i < lhs.size_; ++i)
            if (lhs.values_[i] != rhs.values_[i]) return false;
        return true;
    }

    friend void swap(vector_2d &first, vector_2d &second) noexcept
    {
        std::swap(first.n_row_, second.n_row_);
        std::swap(first.n_col_, second.n_col_);
        std::swap(first.size_, second.size_);
        std::swap(first.values_, second.values_);
    }

  private:
    std::size_t n_row_ = 0;
    std::size_t n_col_ = 0;
    std::size_t size_ = 0;
    std::unique_ptr<T[]> values_;
};
}  // namespace hpxfft::util

#endif  // HPXFFT_GPU_VECTOR_2D_HPP
