// gen_stress.cpp -- regenerates the input stream and the std::map ground truth
// of the reference test tests/large_matrix_stress_tests.cpp:11-90
// (BulkInsertLockSpmvCorrectness): N = 10000, 100000 inserts, mt19937_64(42),
// uniform_int_distribution<size_t>(0, N-1) for indices, <int>(-10, 10) for
// values, zero value = erase, x ~ uniform_real_distribution<double>(-1, 1)
// drawn from the same engine after the inserts.
//
// This is our own program restating the test's data generation (libstdc++'s
// distributions, as the reference's CI uses with gcc); it writes raw
// little-endian arrays to the directory given as argv[1]:
//   ins_row.u64 ins_col.u64 ins_val.f64   -- the insert stream, in order
//   exp_off.u64 exp_col.u64 exp_val.f64   -- ground-truth CSR from std::map
//   x.f64 y_exp.f64                       -- x and the map-order y reference
#include <cstdio>
#include <cstdint>
#include <iterator>
#include <map>
#include <random>
#include <string>
#include <utility>
#include <vector>

template <class T>
static void dump(const std::string &dir, const char *name, const std::vector<T> &v)
{
    std::string p = dir + "/" + name;
    FILE *f = std::fopen(p.c_str(), "wb");
    if (!f) { std::perror(p.c_str()); std::exit(1); }
    if (!v.empty()) std::fwrite(v.data(), sizeof(T), v.size(), f);
    std::fclose(f);
}

int main(int argc, char **argv)
{
    if (argc < 2) { std::fprintf(stderr, "usage: gen_stress OUTDIR\n"); return 2; }
    const std::string dir = argv[1];
    using I = std::size_t;
    using V = double;
    constexpr I N = 10000;
    constexpr int INSERTS = 100000;

    std::mt19937_64 rng(42);
    std::uniform_int_distribution<I> dist_index(0, N - 1);
    std::uniform_int_distribution<int> dist_val(-10, 10);
    std::map<std::pair<I, I>, V> expected;
    std::vector<uint64_t> ins_row, ins_col;
    std::vector<double> ins_val;

    for (int k = 0; k < INSERTS; ++k) {
        I i, j;
        if (k < 10000) {
            i = dist_index(rng);
            j = dist_index(rng);
        } else {
            if (!expected.empty() && (k % 2 == 0)) {
                auto it = expected.begin();
                std::advance(it, static_cast<std::ptrdiff_t>(dist_index(rng) % expected.size()));
                i = it->first.first;
                j = it->first.second;
            } else {
                i = dist_index(rng);
                j = dist_index(rng);
            }
        }
        V val = static_cast<V>(dist_val(rng));
        if (val == 0.0) expected.erase({i, j});
        else expected[{i, j}] = val;
        ins_row.push_back(i);
        ins_col.push_back(j);
        ins_val.push_back(val);
    }

    std::uniform_real_distribution<double> dist_x(-1.0, 1.0);
    std::vector<double> x(N);
    for (I j = 0; j < N; ++j) x[j] = dist_x(rng);

    std::vector<uint64_t> off(N + 1, 0), ecol;
    std::vector<double> eval, y(N, 0.0);
    for (auto const &[key, a] : expected) {
        off[key.first + 1]++;
        ecol.push_back(key.second);
        eval.push_back(a);
        y[key.first] += a * x[key.second];
    }
    for (I r = 0; r < N; ++r) off[r + 1] += off[r];

    dump(dir, "ins_row.u64", ins_row);
    dump(dir, "ins_col.u64", ins_col);
    dump(dir, "ins_val.f64", ins_val);
    dump(dir, "exp_off.u64", off);
    dump(dir, "exp_col.u64", ecol);
    dump(dir, "exp_val.f64", eval);
    dump(dir, "x.f64", x);
    dump(dir, "y_exp.f64", y);
    std::printf("%zu\n", expected.size());
    return 0;
}
