// print_vector.hpp -- hpxfft::util::print_vector_2d with the API of HPX-FFT's
// core/include/hpxfft/util/print_vector.hpp:9-38, on this include directory's
// vector_2d and without HPX.
//
// Prints the buffer as interleaved complex pairs, "(re im) " per pair, one
// line per row and a blank line after the last: the form the reference's
// harness prints a spectrum in (examples/hpxfft/shared_loop_2d.cpp:59-62).
// The reference reads each element into a global `real` that only other
// headers define (print_vector.hpp:22); here the element type is T.
#ifndef HPXFFT_GPU_PRINT_VECTOR_HPP
#define HPXFFT_GPU_PRINT_VECTOR_HPP

#include "vector_2d.hpp"

#include <cstddef>
#include <iostream>

namespace hpxfft::util
{
template <typename T>
void print_vector_2d(const vector_2d<T> &input)
{
    bool opening = true;   // the next element starts a (re im) pair
    for (std::size_t i = 0; i < input.n_row(); ++i)
    {
        for (std::size_t j = 0; j < input.n_col(); ++j)
        {
            if (opening) std::cout << '(' << input(i, j) << ' ';
            else std::cout << input(i, j) << ") ";
            opening = !opening;
        }
        std::cout << '\n';
    }
    std::cout << std::endl;
}
}  // namespace hpxfft::util

#endif
