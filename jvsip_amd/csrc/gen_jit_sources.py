"""Embed the device headers of the run-time compiled kernels into the library.

Usage: python3 gen_jit_sources.py OUT.c HEADER...
Writes a C89 file defining, for each header, its name and its text as a
NUL-terminated byte array, plus the count:
    const int vhip_jit_nheaders;
    const char *const vhip_jit_header_names[];
    const char *const vhip_jit_header_texts[];
The library passes them to hipRTC as in-memory include files.
"""
import os
import sys


def main():
    out, headers = sys.argv[1], sys.argv[2:]
    lines = ["/* generated by gen_jit_sources.py from the device headers; do not edit */"]
    for k, h in enumerate(headers):
        data = open(h, "rb").read() + b"\0"
        lines.append("static const unsigned char vhip_jit_h%d[%d] = {" % (k, len(data)))
        for i in range(0, len(data), 24):
            lines.append("  " + ",".join("%d" % b for b in data[i:i + 24]) + ",")
        lines.append("};")
    lines.append("const int vhip_jit_nheaders = %d;" % len(headers))
    lines.append("const char *const vhip_jit_header_names[] = {%s};"
                 % ", ".join('"%s"' % os.path.basename(h) for h in headers))
    lines.append("const char *const vhip_jit_header_texts[] = {%s};"
                 % ", ".join("(const char *)vhip_jit_h%d" % k for k in range(len(headers))))
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
