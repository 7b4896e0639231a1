// vector_check.cpp -- drives Libo::cuda_abstract_vector<T> (include/cuda_abstract_vector.cuh)
// for T = float, double and int on the GPU and writes every input and result
// as raw little-endian arrays, which tests/test_gpu_vector.py compares with
// numpy.  Each array file holds the x, y and z components back to back.
//
// usage: vector_check <n> <out_dir>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "cuda_abstract_vector.cuh"

using Libo::cuda_abstract_vector;

namespace {

std::string g_dir;

template <typename T>
void dump(const char *name, const cuda_abstract_vector<T> &v) {
  std::vector<T> x, y, z;
  v.export_data(x, y, z);
  const std::string path = g_dir + "/" + name + ".bin";
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  std::fwrite(x.data(), sizeof(T), x.size(), f);
  std::fwrite(y.data(), sizeof(T), y.size(), f);
  std::fwrite(z.data(), sizeof(T), z.size(), f);
  std::fclose(f);
}

void scalar(const char *name, double v) { std::printf("%s %.17g\n", name, v); }

// Deterministic inputs: component c of element i.
template <typename T>
cuda_abstract_vector<T> make(size_t n, double a, double b) {
  std::vector<T> x(n), y(n), z(n);
  for (size_t i = 0; i < n; ++i) {
    x[i] = static_cast<T>(std::sin(a * i + b) * 3.0);
    y[i] = static_cast<T>(std::cos(b * i + a) * 2.0 - 0.25);
    z[i] = static_cast<T>((static_cast<double>(i % 17) - 8.0) * 0.375 + a);
  }
  cuda_abstract_vector<T> v(n);
  v.import_data(x, y, z);
  return v;
}

template <typename T>
void arithmetic(size_t n, const char *tag) {
  const std::string p(tag);
  auto a = make<T>(n, 0.37, 0.11);
  auto b = make<T>(n, 1.3, -0.7);
  dump((p + "_a").c_str(), a);
  dump((p + "_b").c_str(), b);
  dump((p + "_add").c_str(), a + b);
  dump((p + "_sub").c_str(), a - b);
  dump((p + "_neg").c_str(), -a);
  dump((p + "_mul_r").c_str(), a * static_cast<T>(1.75));
  dump((p + "_mul_l").c_str(), static_cast<T>(-0.5) * b);
  cuda_abstract_vector<T> c(a);  // deep copy
  c += b;
  dump((p + "_iadd").c_str(), c);
  c -= a;
  c -= a;
  dump((p + "_isub").c_str(), c);
  c *= static_cast<T>(3.0);
  dump((p + "_imul").c_str(), c);
  cuda_abstract_vector<T> d;
  d = a;  // assignment into an empty vector
  d += d;
  dump((p + "_assign_self_add").c_str(), d);
  dump((p + "_a_after").c_str(), a);  // the copies left a untouched
  scalar((p + "_dot").c_str(), static_cast<double>(a.dot(b)));
  scalar((p + "_norm2").c_str(), static_cast<double>(a.norm2()));
  scalar((p + "_norm").c_str(), static_cast<double>(a.norm()));
  scalar((p + "_dot_fp64").c_str(), a.dot_fp64(b));
  // templated conversion to the other element types
  dump((p + "_to_f32").c_str(), static_cast<cuda_abstract_vector<float>>(a));
  dump((p + "_to_f64").c_str(), static_cast<cuda_abstract_vector<double>>(a));
  dump((p + "_to_i32").c_str(), static_cast<cuda_abstract_vector<int>>(a));
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s <n> <out_dir>\n", argv[0]);
    return 2;
  }
  const size_t n = static_cast<size_t>(std::strtoull(argv[1], nullptr, 10));
  g_dir = argv[2];
  try {
    arithmetic<float>(n, "f32");
    arithmetic<double>(n, "f64");
    // int holds data and converts
    auto k = make<int>(n, 0.9, 0.2);
    dump("i32_a", k);
    cuda_abstract_vector<int> k2(k);
    dump("i32_copy", k2);
    dump("i32_to_f32", static_cast<cuda_abstract_vector<float>>(k));
    dump("i32_to_f64", static_cast<cuda_abstract_vector<double>>(k));
    if (n > 0) {
      k2.add_to_element(n - 1, 5, -6, 7);
      dump("i32_add_elem", k2);
    }
  } catch (const std::exception &e) {
    std::fprintf(stderr, "vector_check: %s\n", e.what());
    return 1;
  }
  std::printf("VECTOR_CHECK_OK %zu\n", n);
  return 0;
}
