"""Writes log_table.hpp: the 128-entry reduction table of table_log
(G_j = 1/(1 + j/128) rounded to double, T_j = -ln G_j correctly rounded),
computed in 60-digit decimal arithmetic.

    python graada3d_amd/csrc/gen_log_table.py
"""
import os
from decimal import Decimal, getcontext

getcontext().prec = 60
N = 128
rows = []
for j in range(N):
    g = 1.0 / (1.0 + j / N)                    # exactly as the device computes it
    t = -Decimal(g).ln()
    rows.append((g, float(t)))
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "log_table.hpp")
with open(out, "w") as f:
    f.write("// Generated by gen_log_table.py: G_j = 1/(1 + j/128), T_j = -ln(G_j).\n")
    f.write("#pragma once\n\n")
    f.write("#if defined(__HIPCC__)\n#define GRA_TABLE_MEM __constant__ const\n#else\n"
            "#define GRA_TABLE_MEM constexpr\n#endif\n\nnamespace gra {\n\n")
    f.write(f"constexpr int kLogTableN = {N};\n")
    f.write("GRA_TABLE_MEM double kLogG[kLogTableN] = {\n")
    f.write(",\n".join(f"    {g!r}" for g, _ in rows) + "};\n")
    f.write("GRA_TABLE_MEM double kLogT[kLogTableN] = {\n")
    f.write(",\n".join(f"    {t!r}" for _, t in rows) + "};\n")
    f.write("\n}  // namespace gra\n")
print(out)
