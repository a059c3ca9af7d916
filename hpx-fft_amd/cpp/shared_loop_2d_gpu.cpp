// shared_loop_2d_gpu -- the harness of HPX-FFT's examples/hpxfft/shared_loop_2d.cpp
// driving hpxfft::shared::loop, here resolved to the MI355X class.
//
// Same command line (shared_loop_2d.cpp:140-146): --nx (default 8) --ny
// (default 14) --plan (estimate) --run (par|seq) --header (false) --result
// (false), given as --key=value or --key value; the two flags take the words
// a program_options bool takes (true/false, 1/0, yes/no, on/off, any case),
// as in the reference's size sweep (benchmark/sbatch_scripts/
// run_hpxfft_size_shared.sh:28, --header=true).  An unknown option or a value
// that does not parse is a usage error: a message and exit status 2.  Same ramp input x[i][j] = j (:33-40),
// same console report (:69-93), same ';'-separated CSV appended to
// runtimes/runtimes_hpx_shared_loop.txt (:95-111), same plan file
// plans/plan_hpx_shared_loop.txt (:114-128).  The n_threads column holds the
// string "gpu".  Sizes: any Nx and even Ny whose lengths the library
// supports (powers of two; lengths within one workgroup's LDS, prime factors
// <= 61 directly and others by Bluestein's algorithm; longer lengths up to
// 2^20 as a product of two such); otherwise the class throws
// std::invalid_argument and the program exits with status 2.
#include "hpxfft/shared/loop.hpp"          // for hpxfft::shared::loop, hpxfft::shared::vector_2d
#include "hpxfft/util/create_dir.hpp"      // for hpxfft::util::create_parent_dir
#include "hpxfft/util/print_vector.hpp"    // for hpxfft::util::print_vector_2d

#include <cctype>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <map>
#include <stdexcept>
#include <string>

// program_options' bool words; throws std::invalid_argument otherwise.
static bool parse_flag(const std::string &key, std::string v)
{
    for (auto &c : v) c = (char)std::tolower((unsigned char)c);
    if (v == "1" || v == "true" || v == "yes" || v == "on") return true;
    if (v == "0" || v == "false" || v == "no" || v == "off") return false;
    throw std::invalid_argument("the argument ('" + v + "') for option '--" + key + "' is invalid");
}

static std::size_t parse_size(const std::string &key, const std::string &v)
{
    std::size_t used = 0;
    unsigned long long x = 0;
    try
    {
        if (v.empty() || v[0] == '-') throw std::invalid_argument(v);
        x = std::stoull(v, &used);
    }
    catch (const std::exception &)
    {
        used = 0;
    }
    if (used == 0 || used != v.size())
        throw std::invalid_argument("the argument ('" + v + "') for option '--" + key + "' is invalid");
    return (std::size_t)x;
}

int main(int argc, char *argv[])
{
    std::map<std::string, std::string> opt = {{"result", "false"}, {"nx", "8"},      {"ny", "14"},
                                              {"plan", "estimate"}, {"run", "par"}, {"header", "false"}};
    bool print_result = false, print_header = false;
    std::size_t dim_c_x = 0, dim_r_y = 0;
    try
    {
        for (int i = 1; i < argc; ++i)
        {
            std::string a = argv[i];
            if (a.rfind("--", 0) != 0) throw std::invalid_argument("unexpected argument '" + a + "'");
            a = a.substr(2);
            const auto eq = a.find('=');
            std::string key = a, val;
            if (eq != std::string::npos)
            {
                key = a.substr(0, eq);
                val = a.substr(eq + 1);
            }
            else if (i + 1 < argc) val = argv[++i];
            else throw std::invalid_argument("option '--" + key + "' requires an argument");
            if (!opt.count(key)) throw std::invalid_argument("unrecognised option '--" + key + "'");
            opt[key] = val;
        }
        print_result = parse_flag("result", opt["result"]);
        print_header = parse_flag("header", opt["header"]);
        dim_c_x = parse_size("nx", opt["nx"]);
        dim_r_y = parse_size("ny", opt["ny"]);
    }
    catch (const std::invalid_argument &e)
    {
        std::cerr << "usage error: " << e.what() << std::endl;
        return 2;
    }
    const std::string run_flag = opt["run"], plan_flag = opt["plan"];
    const std::size_t dim_c_y = dim_r_y / 2 + 1;

    using clock = std::chrono::steady_clock;
    hpxfft::shared::vector_2d values_vec(dim_c_x, 2 * dim_c_y);
    for (std::size_t i = 0; i < dim_c_x; ++i)
        for (std::size_t j = 0; j < dim_r_y; ++j) values_vec(i, j) = (double)j;

    hpxfft::shared::loop fft_computer;
    double total = 0.0, init = 0.0;
    try
    {
        const auto start_total = clock::now();
        fft_computer.initialize(std::move(values_vec), plan_flag);
        const auto stop_init = clock::now();
        if (run_flag == "seq") values_vec = fft_computer.fft_2d_r2c_seq();
        else values_vec = fft_computer.fft_2d_r2c_par();
        const auto stop_total = clock::now();
        total = std::chrono::duration<double>(stop_total - start_total).count();
        init = std::chrono::duration<double>(stop_init - start_total).count();
    }
    catch (const std::invalid_argument &e)
    {
        std::cerr << "invalid argument: " << e.what() << std::endl;
        return 2;
    }
    catch (const std::exception &e)
    {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    if (print_result) hpxfft::util::print_vector_2d(values_vec);

    auto m = [&](const char *k) { return fft_computer.get_measurement(k); };
    std::cout << "\nLocality 0 - shared - " << run_flag << "\n"
              << "Total runtime : " << total << "\n"
              << "Initialization: " << init << "\n"
              << "FFT 2D runtime: " << m("total") << "\n"
              << "FFTW r2c      : " << m("first_fftw") << "\n"
              << "First trans   : " << m("first_trans") << "\n"
              << "FFTW c2c      : " << m("second_fftw") << "\n"
              << "Second trans  : " << m("second_trans") << "\n"
              << "Plan time     : " << m("plan") << "\n"
              << "Plan flops    : " << m("plan_flops") << "\n"
              << std::flush;

    const std::string header = "n_threads;n_x;n_y;plan;run_flag;total;initialization;fft_2d_total;first_fftw;"
                               "first_trans;second_fftw;second_trans;plan_time;plan_flops;\n";
    auto line = [&]() {
        return std::string("gpu;") + std::to_string(dim_c_x) + ";" + std::to_string(dim_r_y) + ";" + plan_flag + ";" +
               run_flag + ";" + std::to_string(total) + ";" + std::to_string(init) + ";" + std::to_string(m("total")) +
               ";" + std::to_string(m("first_fftw")) + ";" + std::to_string(m("first_trans")) + ";" +
               std::to_string(m("second_fftw")) + ";" + std::to_string(m("second_trans")) + ";" +
               std::to_string(m("plan")) + ";" + std::to_string(m("plan_flops")) + ";\n";
    };
    const std::string runtime_file_path = "runtimes/runtimes_hpx_shared_loop.txt";
    hpxfft::util::create_parent_dir(runtime_file_path);
    {
        std::ofstream runtime_file(runtime_file_path, std::ios_base::app);
        if (print_header) runtime_file << header;
        runtime_file << line();
    }
    const std::string plan_file_path = "plans/plan_hpx_shared_loop.txt";
    hpxfft::util::create_parent_dir(plan_file_path);
    {
        std::ofstream plan_info_file(plan_file_path, std::ios_base::app);
        plan_info_file << header << line();
    }
    fft_computer.write_plans_to_file(plan_file_path);
    return 0;
}
