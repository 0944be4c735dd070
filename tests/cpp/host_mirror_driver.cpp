// host_mirror_driver.cpp -- line-oriented driver of the C++23 host mirror
// (include/spira_amd.hpp) for the parity tests: reads one command per line on
// stdin, answers one line per command on stdout.  Values travel as C99 hex
// floats, so every bit (signed zeros included) survives the round trip.
//
//   new serial R C | new parallel R C T      -> ok
//   insert r c v | mark_dirty r | lock | open | clear | rebalance -> ok
//   parallel_fill n r c v ...                -> ok  (parallel_matrix::parallel_fill, the
//                                               triples dealt to their owning partitions)
//   get r c | accumulate r                   -> value <hex>
//   contains r c | is_locked                 -> value 0|1
//   row_nnz r | nnz                          -> value <n>
//   csr                                      -> csr none | csr n nnz ; offs ; cols ; vals
//   spmv serial|parallel|op nx ny x...       -> y <hex>...   (x given, y of size ny)
//   spmv32 serial|parallel nx ny x...        -> y <hex>...   (the float overloads)
//   partition t                              -> part start end ; offs ; cols ; vals
//   bounds                                   -> bounds b...
//   load_csr n nnz offs... cols... vals...   -> ok
//   transpose                                -> ok  (the matrix becomes its transpose:
//                                               serial::algorithms::transpose, or
//                                               parallel::algorithms::transpose)
//   transpose_itself                         -> ok
//   push                                     -> ok  (the matrix moves onto a stack; no current matrix)
//   add | sub | mul | iadd | isub | imul     -> ok  (left = popped stack top, right = the current
//                                               matrix; the result becomes the current matrix:
//                                               MatrixAddition / operator- / spgemm, the parallel
//                                               algorithms when both are parallel_matrix, and the
//                                               in-place operators +=, -=, *= on the left)
// Any exception -> "error <kind> <message>".
#include "spira_amd.hpp"

#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

namespace {

std::unique_ptr<spira_amd::matrix> g_mat;
spira_amd::parallel_matrix *g_par = nullptr;
std::vector<std::unique_ptr<spira_amd::matrix>> g_stack;

std::string hexf(double v)
{
    char buf[64];
    std::snprintf(buf, sizeof buf, "%a", v);
    return buf;
}

double parse_double(const std::string &s) { return std::strtod(s.c_str(), nullptr); }

void print_csr(std::ostream &o, const spira_amd::csr_storage &h)
{
    o << h.n_rows << ' ' << h.nnz << " ;";
    for (auto v : h.offsets) o << ' ' << v;
    o << " ;";
    for (auto v : h.cols) o << ' ' << v;
    o << " ;";
    for (auto v : h.vals) o << ' ' << hexf(v);
}

spira_amd::matrix &mat()
{
    if (!g_mat) throw std::logic_error("driver: no matrix");
    return *g_mat;
}

spira_amd::parallel_matrix &par()
{
    if (!g_par) throw std::logic_error("driver: not a parallel_matrix");
    return *g_par;
}

std::string run(const std::string &line)
{
    std::istringstream in(line);
    std::string cmd;
    in >> cmd;
    std::ostringstream out;
    if (cmd == "new") {
        std::string kind;
        std::size_t r, c;
        in >> kind >> r >> c;
        g_par = nullptr;
        g_mat.reset();
        if (kind == "parallel") {
            std::size_t t;
            in >> t;
            auto p = std::make_unique<spira_amd::parallel_matrix>(r, c, t);
            g_par = p.get();
            g_mat = std::move(p);
        } else {
            g_mat = std::make_unique<spira_amd::matrix>(r, c);
        }
        out << "ok";
    } else if (cmd == "insert") {
        long long r, c;
        std::string v;
        in >> r >> c >> v;
        if (g_par) g_par->insert((uint32_t)r, (uint32_t)c, parse_double(v));
        else mat().insert((uint32_t)r, (uint32_t)c, parse_double(v));
        out << "ok";
    } else if (cmd == "parallel_fill") {
        // parallel_fill n r c v ... : the triples dealt to their owning
        // partitions in the given order, then inserted by parallel_fill.
        std::size_t n;
        in >> n;
        auto &p = par();
        std::vector<std::vector<std::tuple<std::size_t, uint32_t, double>>> by_part(p.n_threads());
        for (std::size_t i = 0; i < n; ++i) {
            long long r, c;
            std::string v;
            in >> r >> c >> v;
            by_part[p.owner((std::size_t)r)].emplace_back((std::size_t)r, (uint32_t)c, parse_double(v));
        }
        p.parallel_fill([&](auto &rows, std::size_t r_start, std::size_t r_end, std::size_t t) {
            for (const auto &[r, c, v] : by_part[t]) {
                if (r < r_start || r >= r_end) throw std::logic_error("driver: triple outside its partition");
                rows[r - r_start].insert(c, v);
            }
        });
        out << "ok";
    } else if (cmd == "mark_dirty") {
        long long r;
        in >> r;
        mat().mark_dirty((uint32_t)r);
        out << "ok";
    } else if (cmd == "lock") {
        if (g_par) g_par->lock();
        else mat().lock();
        out << "ok";
    } else if (cmd == "open") {
        mat().open();
        out << "ok";
    } else if (cmd == "clear") {
        mat().clear();
        out << "ok";
    } else if (cmd == "rebalance") {
        par().rebalance();
        out << "ok";
    } else if (cmd == "get" || cmd == "contains") {
        long long r, c;
        in >> r >> c;
        if (cmd == "get") out << "value " << hexf(mat().get((uint32_t)r, (uint32_t)c));
        else out << "value " << (mat().contains((uint32_t)r, (uint32_t)c) ? 1 : 0);
    } else if (cmd == "accumulate") {
        long long r;
        in >> r;
        out << "value " << hexf(mat().accumulate((uint32_t)r));
    } else if (cmd == "row_nnz") {
        long long r;
        in >> r;
        out << "value " << mat().row_nnz((uint32_t)r);
    } else if (cmd == "nnz") {
        out << "value " << mat().nnz();
    } else if (cmd == "is_locked") {
        out << "value " << (mat().is_locked() ? 1 : 0);
    } else if (cmd == "csr") {
        const auto *h = mat().csr();
        if (!h) {
            out << "csr none";
        } else {
            out << "csr ";
            print_csr(out, *h);
        }
    } else if (cmd == "spmv") {
        std::string kind;
        std::size_t nx, ny;
        in >> kind >> nx >> ny;
        std::vector<double> x(nx);
        for (auto &v : x) {
            std::string s;
            in >> s;
            v = parse_double(s);
        }
        std::vector<double> y(ny);
        if (kind == "parallel") spira_amd::parallel::algorithms::spmv(par(), x, y);
        else if (kind == "op") y = mat() * x;
        else spira_amd::serial::algorithms::spmv(mat(), x, y);
        out << "y";
        for (auto v : y) out << ' ' << hexf(v);
    } else if (cmd == "spmv32") {  // the float overloads: x rounded to FP32, y printed as FP64 hex
        std::string kind;
        std::size_t nx, ny;
        in >> kind >> nx >> ny;
        std::vector<float> x(nx);
        for (auto &v : x) {
            std::string s;
            in >> s;
            v = (float)parse_double(s);
        }
        std::vector<float> y(ny);
        if (kind == "parallel") spira_amd::parallel::algorithms::spmv(par(), x, y);
        else spira_amd::serial::algorithms::spmv(mat(), x, y);
        out << "y";
        for (auto v : y) out << ' ' << hexf((double)v);
    } else if (cmd == "partition") {
        std::size_t t;
        in >> t;
        const auto p = par().partition_at(t);
        out << "part " << p.row_start << ' ' << p.row_end << " ; ";
        print_csr(out, p.csr);
    } else if (cmd == "bounds") {
        out << "bounds";
        for (auto b : par().partition_bounds()) out << ' ' << b;
    } else if (cmd == "load_csr") {
        spira_amd::csr_storage h;
        std::size_t n, nnz;
        in >> n >> nnz;
        h.n_rows = n;
        h.nnz = nnz;
        h.offsets.resize(n + 1);
        for (auto &v : h.offsets) in >> v;
        h.cols.resize(nnz);
        for (auto &v : h.cols) in >> v;
        h.vals.resize(nnz);
        for (auto &v : h.vals) {
            std::string s;
            in >> s;
            v = parse_double(s);
        }
        mat().load_csr(std::move(h));
        out << "ok";
    } else if (cmd == "transpose") {
        if (g_par) {
            auto p = std::make_unique<spira_amd::parallel_matrix>(spira_amd::parallel::algorithms::transpose(*g_par));
            g_par = p.get();
            g_mat = std::move(p);
        } else {
            g_mat = std::make_unique<spira_amd::matrix>(spira_amd::serial::algorithms::transpose(mat()));
        }
        out << "ok";
    } else if (cmd == "push") {
        g_stack.push_back(std::move(g_mat));
        g_par = nullptr;
        out << "ok";
    } else if (cmd == "add" || cmd == "sub" || cmd == "mul" || cmd == "iadd" || cmd == "isub" || cmd == "imul") {
        if (g_stack.empty() || !g_stack.back()) throw std::logic_error("driver: empty stack");
        std::unique_ptr<spira_amd::matrix> left = std::move(g_stack.back());
        g_stack.pop_back();
        auto *lp = dynamic_cast<spira_amd::parallel_matrix *>(left.get());
        if (lp && g_par && (cmd == "add" || cmd == "mul")) {
            auto p = std::make_unique<spira_amd::parallel_matrix>(
                cmd == "add" ? spira_amd::parallel::algorithms::MatrixAddition(*lp, *g_par)
                             : spira_amd::parallel::algorithms::spgemm(*lp, *g_par));
            g_par = p.get();
            g_mat = std::move(p);
        } else {
            const spira_amd::matrix &right = mat();
            std::unique_ptr<spira_amd::matrix> res;
            if (cmd == "add") res = std::make_unique<spira_amd::matrix>(spira_amd::serial::algorithms::MatrixAddition(*left, right));
            else if (cmd == "sub") res = std::make_unique<spira_amd::matrix>(*left - right);
            else if (cmd == "mul") res = std::make_unique<spira_amd::matrix>(spira_amd::serial::algorithms::spgemm(*left, right));
            else {
                spira_amd::matrix l2 = std::move(*left);
                if (cmd == "iadd") l2 += right;
                else if (cmd == "isub") l2 -= right;
                else l2 *= right;
                res = std::make_unique<spira_amd::matrix>(std::move(l2));
            }
            g_par = nullptr;
            g_mat = std::move(res);
        }
        out << "ok";
    } else if (cmd == "transpose_itself") {
        spira_amd::serial::algorithms::transpose_itself(mat());
        out << "ok";
    } else {
        out << "error driver unknown command " << cmd;
    }
    return out.str();
}

}  // namespace

int main()
{
    std::ios::sync_with_stdio(false);
    std::string line;
    while (std::getline(std::cin, line)) {
        if (line.empty()) continue;
        if (line == "quit") break;
        std::string reply;
        try {
            reply = run(line);
        } catch (const std::invalid_argument &e) {
            reply = std::string("error invalid_argument ") + e.what();
        } catch (const std::out_of_range &e) {
            reply = std::string("error out_of_range ") + e.what();
        } catch (const std::length_error &e) {
            reply = std::string("error length_error ") + e.what();
        } catch (const std::logic_error &e) {
            reply = std::string("error logic_error ") + e.what();
        } catch (const std::bad_alloc &) {
            reply = "error bad_alloc";
        } catch (const std::exception &e) {
            reply = std::string("error runtime_error ") + e.what();
        }
        std::cout << reply << '\n' << std::flush;
    }
    g_par = nullptr;
    g_mat.reset();
    return 0;
}
