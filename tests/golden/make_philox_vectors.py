"""Writes philox_rocrand.json: Philox4x32-10 outputs computed by ROCm's
rocRAND implementation (/opt/rocm/include/rocrand/rocrand_philox4x32_10.h,
open source), compiled for the host.  The oracle's own Philox is checked
against these vectors and the Random123 known answers.

    python tests/golden/make_philox_vectors.py
"""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = r'''
#include <cstdio>
#include <cstdint>
#include <rocrand/rocrand_philox4x32_10.h>
struct Probe : rocrand_device::philox4x32_10_engine {
  Probe() : rocrand_device::philox4x32_10_engine(0, 0, 0) {}
  uint4 run(uint4 c, uint2 k) { return ten_rounds(c, k); }
};
int main() {
  Probe p;
  uint64_t s = 0x9E3779B97F4A7C15ull;
  auto next = [&]() { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return (unsigned)(s >> 11); };
  std::printf("[");
  for (int i = 0; i < 64; ++i) {
    uint4 c; uint2 k;
    if (i == 0) { c = {0,0,0,0}; k = {0,0}; }
    else if (i == 1) { c = {0xffffffffu,0xffffffffu,0xffffffffu,0xffffffffu}; k = {0xffffffffu,0xffffffffu}; }
    else { c = {next(), next(), next(), next()}; k = {next(), next()}; }
    uint4 o = p.run(c, k);
    std::printf("%s[[%u,%u,%u,%u],[%u,%u],[%u,%u,%u,%u]]", i ? "," : "", c.x,c.y,c.z,c.w,k.x,k.y,o.x,o.y,o.z,o.w);
  }
  std::printf("]\n");
}
'''

def main():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "p.cpp")
        exe = os.path.join(d, "p")
        open(src, "w").write(SRC)
        subprocess.run(["/opt/rocm/bin/hipcc", "-O1", "-o", exe, src], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    rows = json.loads(out)
    vectors = [{"ctr": r[0], "key": r[1], "out": r[2]} for r in rows]
    with open(os.path.join(HERE, "philox_rocrand.json"), "w") as f:
        f.write('{"source": "rocrand_philox4x32_10.h ten_rounds (ROCm 7.2, host build)",\n')
        f.write(' "vectors": [\n')
        f.write(",\n".join("  " + json.dumps(v) for v in vectors))
        f.write("\n]}\n")
    print(f"wrote {len(vectors)} vectors")

if __name__ == "__main__":
    main()
