// test_vector_2d -- the reference's container tests (test/src/test_vector_2d.cpp:6-33)
// against include/hpxfft/util/vector_2d.hpp, plus the RAII behaviour this
// version adds (copies are deep, moved-from objects are empty).  CPU only.
#include "hpxfft/util/vector_2d.hpp"

#include <iostream>
#include <stdexcept>
#include <utility>

static int failures = 0;
#define CHECK(cond)                                                                                                    \
    do                                                                                                                 \
    {                                                                                                                  \
        if (!(cond))                                                                                                   \
        {                                                                                                              \
            std::cerr << "FAILED: " #cond " (line " << __LINE__ << ")\n";                                             \
            ++failures;                                                                                                \
        }                                                                                                              \
    } while (0)

int main()
{
    using hpxfft::util::vector_2d;
    {
        vector_2d<double> vec(3, 3, 4.0);   // test_vector_2d.cpp:6-16
        CHECK(vec.n_row() == 3);
        CHECK(vec.n_col() == 3);
        CHECK(vec(0, 0) == 4.0);
        CHECK(vec(2, 2) == 4.0);
        CHECK(vec(1, 2) == 4.0);
        CHECK(vec.size() == 9);
    }
    {
        vector_2d<double> vec(3, 3, 1.0);   // :18-23
        bool threw = false;
        try
        {
            (void)vec.at(3, 3);
        }
        catch (const std::runtime_error &)
        {
            threw = true;
        }
        CHECK(threw);
    }
    {
        vector_2d<double> vec1(2, 2, 5.0), vec2(2, 2, 5.0), vec3(2, 2, 6.0);   // :25-33
        CHECK(vec1 == vec2);
        CHECK(!(vec1 == vec3));
    }
    {
        vector_2d<double> z(2, 3);   // zero fill
        for (double v : z) CHECK(v == 0.0);
        CHECK(z.row(1) == z.data() + 3);
        vector_2d<double> a(2, 2, 1.0);
        vector_2d<double> b = a;   // deep copy
        b(0, 0) = 7.0;
        CHECK(a(0, 0) == 1.0);
        vector_2d<double> c = std::move(b);
        CHECK(b.size() == 0 && b.data() == nullptr);
        CHECK(c(0, 0) == 7.0);
        a = c;
        CHECK(a == c && c(0, 0) == 7.0);
    }
    std::cout << (failures ? "FAILED" : "ALL PASSED") << std::endl;
    return failures ? 1 : 0;
}
