// GSD 2.0 trajectory writer of the drop-in front end (the file format of
// PS_Box::writeGSDtraj, PS_Box_gsdUtils.cu:14-160).  Layout on disk:
//
//   256-byte header | chunk data of every frame | index | name list
//
// The index (32-byte entries sorted by frame, then by name id, unused entries
// zero) and the name list (null-terminated names, zero padded to a multiple of
// 64 bytes) sit after the data and are rewritten, with the header, at the end
// of every frame, so the file is complete after each frame; the next frame's
// data overwrites them before they are written again further on.
#include <algorithm>
#include <cstring>
#include <fcntl.h>
#include <unistd.h>

#include "matilda.h"

namespace {

struct GsdHeader {
    uint64_t magic;
    uint64_t index_location;
    uint64_t index_allocated_entries;
    uint64_t namelist_location;
    uint64_t namelist_allocated_entries;
    uint32_t schema_version;
    uint32_t gsd_version;
    char application[64];
    char schema[64];
    char reserved[80];
};
static_assert(sizeof(GsdHeader) == 256, "GSD header is 256 bytes");

struct GsdEntry {
    uint64_t frame;
    uint64_t N;
    int64_t location;
    uint32_t M;
    uint16_t id;
    uint8_t type;
    uint8_t flags;
};
static_assert(sizeof(GsdEntry) == 32, "GSD index entry is 32 bytes");

size_t type_size(uint8_t t)
{
    switch (t) {
    case GsdWriter::UINT8: case GsdWriter::INT8: return 1;
    case GsdWriter::UINT32: case GsdWriter::INT32: case GsdWriter::FLOAT: return 4;
    case GsdWriter::UINT64: return 8;
    default: return 0;
    }
}

void write_all(int fd, const void* p, size_t n, uint64_t off, const std::string& path)
{
    const char* c = (const char*)p;
    while (n > 0) {
        ssize_t w = pwrite(fd, c, n, (off_t)off);
        if (w <= 0) die("GSD: write to " + path + " failed");
        c += w;
        n -= (size_t)w;
        off += (uint64_t)w;
    }
}

} // namespace

GsdWriter::GsdWriter(const std::string& path, const std::string& application, const std::string& schema,
                     uint32_t schema_version)
    : path_(path), application_(application), schema_(schema), schema_version_(schema_version)
{
    fd_ = open(path.c_str(), O_RDWR | O_CREAT | O_TRUNC, 0644);
    if (fd_ < 0) die("GSD: cannot create " + path);
    data_end_ = sizeof(GsdHeader);
    flush();
}

GsdWriter::~GsdWriter()
{
    if (fd_ >= 0) close(fd_);
}

void GsdWriter::chunk(const std::string& name, Type type, uint64_t N, uint32_t M, const void* data)
{
    const size_t ts = type_size((uint8_t)type);
    if (ts == 0 || M == 0 || (N > 0 && !data) || name.empty() || name.size() >= 64)
        die("GSD: invalid chunk " + name);
    uint16_t id;
    auto it = ids_.find(name);
    if (it == ids_.end()) {
        id = (uint16_t)names_.size();
        ids_[name] = id;
        names_.push_back(name);
    } else {
        id = it->second;
    }
    const uint64_t bytes = N * M * ts;
    if (bytes) write_all(fd_, data, bytes, data_end_, path_);
    pending_.push_back(Entry{frame_, N, (int64_t)data_end_, M, id, (uint8_t)type});
    data_end_ += bytes;
}

void GsdWriter::end_frame()
{
    std::stable_sort(pending_.begin(), pending_.end(), [](const Entry& a, const Entry& b) { return a.id < b.id; });
    entries_.insert(entries_.end(), pending_.begin(), pending_.end());
    pending_.clear();
    ++frame_;
    flush();
}

void GsdWriter::flush()
{
    // index: every entry's frame must stay below the allocated entry count
    uint64_t alloc = 128;
    while (alloc < entries_.size() + 1 || alloc <= frame_) alloc *= 2;
    std::vector<GsdEntry> index(alloc);
    std::memset(index.data(), 0, sizeof(GsdEntry) * alloc);
    for (size_t k = 0; k < entries_.size(); ++k) {
        const Entry& e = entries_[k];
        index[k] = GsdEntry{e.frame, e.N, e.location, e.M, e.id, e.type, 0};
    }
    std::string names;
    for (const auto& n : names_) {
        names += n;
        names.push_back('\0');
    }
    names.push_back('\0');
    names.resize((names.size() + 63) / 64 * 64, '\0');
    GsdHeader h;
    std::memset(&h, 0, sizeof(h));
    h.magic = 0x65DF65DF65DF65DFull;
    h.index_location = data_end_;
    h.index_allocated_entries = alloc;
    h.namelist_location = data_end_ + sizeof(GsdEntry) * alloc;
    h.namelist_allocated_entries = names.size() / 64;
    h.schema_version = schema_version_;
    h.gsd_version = 2u << 16;
    std::strncpy(h.application, application_.c_str(), sizeof(h.application) - 1);
    std::strncpy(h.schema, schema_.c_str(), sizeof(h.schema) - 1);
    write_all(fd_, index.data(), sizeof(GsdEntry) * alloc, h.index_location, path_);
    write_all(fd_, names.data(), names.size(), h.namelist_location, path_);
    write_all(fd_, &h, sizeof(h), 0, path_);
    if (ftruncate(fd_, (off_t)(h.namelist_location + names.size())) != 0) die("GSD: truncating " + path_ + " failed");
}
