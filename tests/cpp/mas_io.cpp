// Host-only driver of include/psum/mas.hpp for the format tests.
//   mas_io write <file>   writes a fixed set of blocks, masks one, replaces one
//   mas_io read <file>    prints "name type shape... | sum of elements" per block
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "psum/mas.hpp"

using psum::serialization::mas_file;

int main(int argc, char** argv)
{
    if (argc != 3) return 2;
    const std::string mode = argv[1], path = argv[2];
    try {
        if (mode == "write") {
            mas_file f(path, mas_file::replaceMode);
            const double d[6] = {1.5, -2.0, 3.25, 4.0, 5.5, -6.75};
            f.writeData<double>("grid.lower", d, {2, 3});
            const float s[2] = {0.5f, -1.25f};
            f.writeData<float>("single", s, {2, 1});
            const int32_t c[3] = {-7, 0, 2147483647};
            f.writeData<int32_t>("cells", c, {3, 1});
            const uint64_t u[1] = {18446744073709551615ull};
            f.writeData<uint64_t>("big", u, {1, 1});
            const bool b[2] = {true, false};
            f.writeData<bool>("flags", b, {2, 1});
            f.writeData("label", std::string("[#0]MI355X"));
            f.writeScalar<int64_t>("step", -42);
            const int8_t m[2] = {-128, 127};
            f.writeData<int8_t>("masked", m, {2, 1});
            f.smashData("masked");
            const double r[1] = {9.0};
            f.writeData<double>("replaced", r, {1, 1});
            const double r2[2] = {10.0, 11.0};
            f.replaceData<double>("replaced", r2, {2, 1});
            bool dup = false;
            try {
                f.writeData<double>("step", r, {1, 1});
            } catch (const psum::serialization::mas_duplicate_block_error&) {
                dup = true;
            }
            if (!dup) return 3;
            return 0;
        }
        if (mode == "read") {
            mas_file f(path);
            for (const auto& h : f.getHeads()) {
                std::printf("%s %s", h.info.blockName.c_str(), h.info.blockType.c_str());
                for (size_t s : h.info.blockShape) std::printf(" %zu", s);
                const auto b = f.readData(h.info.blockName);
                double sum = 0;
                const std::string& t = h.info.blockType;
                for (size_t i = 0; i < b.info.blockSize; ++i) {
                    if (t == "double") sum += b.data<double>()[i];
                    else if (t == "single") sum += b.data<float>()[i];
                    else if (t == "int64") sum += (double)b.data<int64_t>()[i];
                    else if (t == "uint64") sum += (double)b.data<uint64_t>()[i];
                    else if (t == "int32") sum += b.data<int32_t>()[i];
                    else if (t == "int8") sum += b.data<int8_t>()[i];
                    else if (t == "uint8") sum += b.data<uint8_t>()[i];
                    else if (t == "logical") sum += b.data<bool>()[i];
                    else if (t == "char") sum += (unsigned char)b.data<char>()[i];
                }
                std::printf(" | %.17g\n", sum);
            }
            return 0;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 2;
}
