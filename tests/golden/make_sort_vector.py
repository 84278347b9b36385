"""Writes tests/golden/counting_sort_n32_seed123.json.

The reference's only unit test of its GPU path (ot/cuda/test.cu:17-84) sorts
n=32 values drawn by std::default_random_engine(123) with
std::uniform_int_distribution<int>(1, 31) and asserts
inputs[i] == outputs[orders[i]] (test.cu:73-78).  This script reproduces those
inputs with the same libstdc++ generator (compiled here with g++) and stores
them with the sorted outputs the counting sort must produce.
"""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = r"""
#include <cstdio>
#include <random>
int main() {
  const int n = 32, amin = 1, amax = n - 1;
  std::default_random_engine gen(123);
  std::uniform_int_distribution<int> dist(amin, amax);
  for (int i = 0; i < n; ++i) { int v = dist(gen); std::printf("%d\n", v < amin ? amin : v); }
  return 0;
}
"""


def main():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "gen.cpp")
        exe = os.path.join(d, "gen")
        with open(src, "w") as f:
            f.write(SRC)
        subprocess.run(["g++", "-O2", "-o", exe, src], check=True)
        vals = [int(x) for x in subprocess.run([exe], capture_output=True, text=True,
                                               check=True).stdout.split()]
    out = {"n": 32, "amin": 1, "amax": 31, "inputs": vals, "sorted": sorted(vals),
           "source": "ot/cuda/test.cu:17-84 (default_random_engine(123), U{1..31})"}
    with open(os.path.join(HERE, "counting_sort_n32_seed123.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out)


if __name__ == "__main__":
    main()
