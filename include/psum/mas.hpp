// MAS (multi-array set) files, the reference's checkpoint and fixture format
// (src/serialization/mas.hpp:49-453, foundation.hpp:75-113).  A file is a
// sequence of blocks
//
//     "#MAS!#"  int32 key_len  int32 type_len  int32 shape_len
//     key  type  size_t[shape_len] shape  data (prod(shape) elements)
//
// in native byte order, with MATLAB type names (double, single, int8 ..
// uint64, logical, char).  Deleting a block ("smash") overwrites the first
// byte of its key with '!' in place; scans skip such blocks.  Keys may not be
// empty or begin with '!', and a key may be written once (replaceData smashes
// the old block first).
//
// Header-only, C++17, host code; the object overloads for grids, fields and
// particle groups are in serialization.hpp.
#pragma once

#include <cstdint>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

namespace psum {
namespace serialization {

struct mas_block_info {
    std::string blockName;
    std::string blockType;
    std::vector<size_t> blockShape;
    size_t blockSize = 0;   // elements
};

struct mas_block_head {
    mas_block_info info;
    size_t position = 0;   // byte offset of the block's magic string
};

// a block read back: its header and raw bytes (empty name: not found)
struct mas_block {
    mas_block_info info;
    std::vector<unsigned char> bytes;
    template <typename T>
    const T* data() const
    {
        return reinterpret_cast<const T*>(bytes.data());
    }
};

class mas_duplicate_block_error : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

// MATLAB name and element size of a block type
template <typename T>
constexpr const char* mas_type_name()
{
    if constexpr (std::is_same_v<T, double>) return "double";
    else if constexpr (std::is_same_v<T, float>) return "single";
    else if constexpr (std::is_same_v<T, bool>) return "logical";
    else if constexpr (std::is_same_v<T, char>) return "char";
    else if constexpr (std::is_integral_v<T> && std::is_signed_v<T> && sizeof(T) == 8) return "int64";
    else if constexpr (std::is_integral_v<T> && std::is_unsigned_v<T> && sizeof(T) == 8) return "uint64";
    else if constexpr (std::is_integral_v<T> && std::is_signed_v<T> && sizeof(T) == 4) return "int32";
    else if constexpr (std::is_integral_v<T> && std::is_unsigned_v<T> && sizeof(T) == 4) return "uint32";
    else if constexpr (std::is_integral_v<T> && std::is_signed_v<T> && sizeof(T) == 2) return "int16";
    else if constexpr (std::is_integral_v<T> && std::is_unsigned_v<T> && sizeof(T) == 2) return "uint16";
    else if constexpr (std::is_integral_v<T> && std::is_signed_v<T> && sizeof(T) == 1) return "int8";
    else if constexpr (std::is_integral_v<T> && std::is_unsigned_v<T> && sizeof(T) == 1) return "uint8";
    else static_assert(sizeof(T) == 0, "MAS: unsupported element type");
}

inline size_t mas_type_size(const std::string& t)
{
    static const std::unordered_map<std::string, size_t> sz{
        {"double", 8}, {"single", 4}, {"int64", 8}, {"uint64", 8}, {"int32", 4}, {"uint32", 4},
        {"int16", 2},  {"uint16", 2}, {"int8", 1},  {"uint8", 1},  {"logical", 1}, {"char", 1}};
    auto it = sz.find(t);
    if (it == sz.end()) throw std::runtime_error("MAS: unknown block type '" + t + "'");
    return it->second;
}

class mas_file {
public:
    // autoMode: open (create when absent) and scan; replaceMode: start an
    // empty file; quickMode: open an existing file without scanning
    enum OpenMode { autoMode, replaceMode, quickMode };

    explicit mas_file(const std::string& path, OpenMode mode = autoMode) : filename_(path)
    {
        std::ifstream probe(path, std::ios::binary);
        const bool exists = probe.good();
        probe.close();
        if (!exists && mode == quickMode) throw std::runtime_error("MAS: file not found: " + path);
        if (!exists) {
            const std::filesystem::path parent = std::filesystem::path(path).parent_path();
            if (!parent.empty()) std::filesystem::create_directories(parent);
        }
        if (!exists || mode == replaceMode) clear();
        else if (mode == autoMode) readHead();
    }

    void clear()
    {
        std::ofstream ofs(filename_, std::ios::binary | std::ios::trunc);
        if (!ofs) throw std::runtime_error("MAS: cannot create " + filename_);
        heads_.clear();
        index_.clear();
    }

    template <typename T>
    void writeData(const std::string& key, const T* data, const std::vector<size_t>& shape)
    {
        if (key.empty()) throw std::invalid_argument("MAS: key can't be empty");
        if (key[0] == '!') throw std::invalid_argument("MAS: key can't start with '!'");
        if (index_.count(key)) throw mas_duplicate_block_error("MAS: block already exists: " + key);
        const std::string type = mas_type_name<T>();
        std::ofstream ofs(filename_, std::ios::binary | std::ios::app);
        if (!ofs) throw std::runtime_error("MAS: cannot open " + filename_);
        ofs.seekp(0, std::ios::end);
        const size_t position = (size_t)ofs.tellp();
        const int32_t lens[3] = {(int32_t)key.size(), (int32_t)type.size(), (int32_t)shape.size()};
        ofs.write(kMagic, 6);
        ofs.write(reinterpret_cast<const char*>(lens), sizeof(lens));
        ofs.write(key.data(), (std::streamsize)key.size());
        ofs.write(type.data(), (std::streamsize)type.size());
        size_t count = 1;
        for (size_t s : shape) {
            ofs.write(reinterpret_cast<const char*>(&s), sizeof(size_t));
            count *= s;
        }
        ofs.write(reinterpret_cast<const char*>(data), (std::streamsize)(count * sizeof(T)));
        if (!ofs) throw std::runtime_error("MAS: write failed: " + filename_);
        heads_.push_back({{key, type, shape, count}, position});
        index_[key] = heads_.size() - 1;
    }
    void writeData(const std::string& key, const std::string& s)
    {
        writeData<char>(key, s.data(), {s.size(), 1});
    }
    template <typename T, typename = std::enable_if_t<std::is_arithmetic_v<T>>>
    void writeScalar(const std::string& key, T v)
    {
        writeData<T>(key, &v, {1, 1});
    }

    // scan every block header; masked blocks are skipped
    std::vector<mas_block_head>& readHead()
    {
        heads_.clear();
        index_.clear();
        std::ifstream ifs(filename_, std::ios::binary);
        if (!ifs) throw std::runtime_error("MAS: cannot open " + filename_);
        ifs.seekg(0, std::ios::end);
        const size_t size = (size_t)ifs.tellg();
        size_t pos = 0;
        while (pos < size) {
            ifs.seekg((std::streamoff)pos);
            char magic[6];
            int32_t lens[3];
            ifs.read(magic, 6);
            ifs.read(reinterpret_cast<char*>(lens), sizeof(lens));
            if (!ifs || std::memcmp(magic, kMagic, 6) != 0 || lens[0] < 0 || lens[1] < 0 || lens[2] < 0)
                throw std::runtime_error("MAS: bad block header at byte " + std::to_string(pos) + " of " + filename_);
            mas_block_info info;
            info.blockName.resize((size_t)lens[0]);
            info.blockType.resize((size_t)lens[1]);
            ifs.read(&info.blockName[0], lens[0]);
            ifs.read(&info.blockType[0], lens[1]);
            info.blockShape.resize((size_t)lens[2]);
            size_t count = 1;
            for (auto& s : info.blockShape) {
                ifs.read(reinterpret_cast<char*>(&s), sizeof(size_t));
                count *= s;
            }
            if (!ifs) throw std::runtime_error("MAS: truncated block header in " + filename_);
            info.blockSize = count;
            const size_t data_at = pos + 6 + sizeof(lens) + (size_t)lens[0] + (size_t)lens[1] + 8 * (size_t)lens[2];
            const size_t end = data_at + count * mas_type_size(info.blockType);
            if (end > size) throw std::runtime_error("MAS: incomplete block '" + info.blockName + "'");
            if (!info.blockName.empty() && info.blockName[0] != '!') {
                heads_.push_back({info, pos});
                index_[info.blockName] = heads_.size() - 1;
            }
            pos = end;
        }
        return heads_;
    }

    // mask the block(s) of that name (and type, unless "*"); true if any
    bool smashData(const std::string& key, const std::string& type = "*")
    {
        bool found = false;
        std::fstream fs(filename_, std::ios::in | std::ios::out | std::ios::binary);
        for (const auto& h : heads_)
            if (h.info.blockName == key && (type == "*" || h.info.blockType == type)) {
                fs.seekp((std::streamoff)(h.position + 6 + 3 * sizeof(int32_t)));
                fs.put('!');
                found = true;
            }
        fs.close();
        if (found) readHead();
        return found;
    }

    template <typename T>
    void replaceData(const std::string& key, const T* data, const std::vector<size_t>& shape)
    {
        smashData(key);
        writeData<T>(key, data, shape);
    }

    mas_block readData(const std::string& key, const std::string& type = "*")
    {
        if (!index_.count(key)) readHead();
        auto it = index_.find(key);
        if (it == index_.end()) return {};
        const mas_block_head& h = heads_[it->second];
        if (type != "*" && h.info.blockType != type) return {};
        mas_block b;
        b.info = h.info;
        b.bytes.resize(h.info.blockSize * mas_type_size(h.info.blockType));
        std::ifstream ifs(filename_, std::ios::binary);
        ifs.seekg((std::streamoff)(h.position + 6 + 3 * sizeof(int32_t) + h.info.blockName.size() +
                                   h.info.blockType.size() + 8 * h.info.blockShape.size()));
        ifs.read(reinterpret_cast<char*>(b.bytes.data()), (std::streamsize)b.bytes.size());
        if (!ifs) throw std::runtime_error("MAS: incomplete block '" + key + "'");
        return b;
    }
    mas_block readData(int index)
    {
        if (index < 0 || (size_t)index >= heads_.size()) throw std::out_of_range("MAS: block index out of range");
        return readData(heads_[(size_t)index].info.blockName);
    }

    // typed read of a whole block; the type must match exactly
    template <typename T>
    std::vector<T> read(const std::string& key, std::vector<size_t>* shape = nullptr)
    {
        mas_block b = readData(key);
        if (b.info.blockName.empty()) throw std::runtime_error("MAS: no block '" + key + "' in " + filename_);
        if (b.info.blockType != mas_type_name<T>())
            throw std::runtime_error("MAS: block '" + key + "' has type " + b.info.blockType + ", not " +
                                     mas_type_name<T>());
        if (shape) *shape = b.info.blockShape;
        std::vector<T> out(b.info.blockSize);
        if (!out.empty()) std::memcpy(out.data(), b.bytes.data(), b.bytes.size());
        return out;
    }

    size_t blockNum() const { return heads_.size(); }
    const std::vector<mas_block_head>& getHeads() const { return heads_; }
    const std::string& getFilename() const { return filename_; }

private:
    static constexpr const char* kMagic = "#MAS!#";
    std::string filename_;
    std::vector<mas_block_head> heads_;
    std::unordered_map<std::string, size_t> index_;
};

}  // namespace serialization
}  // namespace psum
