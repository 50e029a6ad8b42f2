// The reference's on-disk web dump for pageRank, read into (and written from) an out-link CSR.
//
// Format (readWebMetaData / readWebPagesFile / mapAllFiles, testSuite/pageRank/source/code/src/
// pageRank.cpp:53-197, and the writer testSuite/pageRank/scripts/webGenerator.pl:121-274):
//   <dir>/web_dump_metadata          three uint32: pages, max out-links per page, pages per file
//   <dir>/web/page_<1 + i * per_file> the pages i * per_file + 1 .. of file i (the last file may
//                                    hold fewer), each a fixed record of 1 + max_out uint32:
//                                    the out-link count, that many 1-based target page ids, then
//                                    zero padding up to max_out.
// Duplicate targets and self links are legal and kept (each link counts, as in the reference's
// accumulation). The files are mapped with pmMapFile (the reference maps them the same way, then
// streams the padded records to the device on every iteration); here they are parsed once, in
// parallel over files, into the out-link CSR that ua_pr_create turns into the resident graph.
#include "app_common.h"
#include "unicorn_apps.h"

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <string>
#include <sys/stat.h>
#include <thread>
#include <vector>

namespace {

struct Meta {
    uint32_t pages = 0, max_out = 0, per_file = 0;
};

std::string page_file(const std::string& dir, uint64_t first_page0) {
    return dir + "/web/page_" + std::to_string(first_page0 + 1);
}

int fail(const std::string& msg) {
    apps::set_error(msg);
    return -1;
}

int read_meta(const std::string& dir, Meta* m) {
    const std::string path = dir + "/web_dump_metadata";
    FILE* fp = fopen(path.c_str(), "rb");
    if (!fp) return fail("web dump: cannot open " + path);
    uint32_t v[3];
    const size_t got = fread(v, 4, 3, fp);
    fclose(fp);
    if (got != 3) return fail("web dump: short metadata file " + path);
    m->pages = v[0];
    m->max_out = v[1];
    m->per_file = v[2];
    if (!m->pages || !m->per_file) return fail("web dump: metadata with zero pages or pages per file");
    return 0;
}

// The mapped records of file f (pages f * per_file .. ), checked against the file's size.
int map_file(const std::string& dir, const Meta& m, uint64_t f, const uint32_t** recs, uint32_t* npages,
             std::string* err) {
    const uint64_t p0 = f * m.per_file;
    const uint64_t n = std::min<uint64_t>(m.per_file, m.pages - p0);
    const std::string path = page_file(dir, p0);
    struct stat sb;
    if (stat(path.c_str(), &sb) != 0) { *err = "web dump: cannot open " + path; return -1; }
    const uint64_t want = n * ((uint64_t)m.max_out + 1) * 4;
    if ((uint64_t)sb.st_size < want) { *err = "web dump: " + path + " is shorter than its records"; return -1; }
    if (pm::pmMapFile(path.c_str()) != pm::pmSuccess) { *err = "web dump: cannot map " + path; return -1; }
    *recs = (const uint32_t*)pm::pmGetMappedFile(path.c_str());
    *npages = (uint32_t)n;
    if (!*recs) { *err = "web dump: mapping of " + path + " failed"; return -1; }
    return 0;
}

void unmap_files(const std::string& dir, const Meta& m, uint64_t files) {
    for (uint64_t f = 0; f < files; ++f) pm::pmUnmapFile(page_file(dir, f * m.per_file).c_str());
}

// Runs body(f) for every file on up to 16 threads; the first error wins.
template <typename F>
int for_files(uint64_t files, F body) {
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const unsigned nt = (unsigned)std::min<uint64_t>(hw, files);
    std::atomic<uint64_t> next{0};
    std::atomic<int> rc{0};
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&] {
            for (uint64_t f; rc.load() == 0 && (f = next.fetch_add(1)) < files;)
                if (body(f) != 0) rc.store(-1);
        });
    for (auto& x : th) x.join();
    return rc.load();
}

}  // namespace

extern "C" int ua_webdump_meta(const char* dir, uint32_t* pages, uint32_t* max_out, uint32_t* per_file) {
    if (!dir || !pages || !max_out || !per_file) return fail("ua_webdump_meta: null argument");
    Meta m;
    if (read_meta(dir, &m)) return -1;
    *pages = m.pages;
    *max_out = m.max_out;
    *per_file = m.per_file;
    return 0;
}

// Phase 1 (col == NULL): row_ptr[0..pages] and *edges. Phase 2: also col (0-based targets,
// capacity col_capacity), sources in page order and targets in record order.
extern "C" int ua_webdump_read(const char* dir_, uint32_t* row_ptr, uint32_t* col, uint64_t col_capacity,
                               uint64_t* edges) {
    if (!dir_ || !row_ptr || !edges) return fail("ua_webdump_read: null argument");
    const std::string dir = dir_;
    Meta m;
    if (read_meta(dir, &m)) return -1;
    const uint64_t files = (m.pages + (uint64_t)m.per_file - 1) / m.per_file;
    const uint64_t stride = (uint64_t)m.max_out + 1;
    std::vector<std::string> errors(files);
    // counts into row_ptr[p + 1]
    int rc = for_files(files, [&](uint64_t f) {
        const uint32_t* r;
        uint32_t n;
        if (map_file(dir, m, f, &r, &n, &errors[f])) return -1;
        const uint64_t p0 = f * m.per_file;
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t c = r[i * stride];
            if (c > m.max_out) {
                errors[f] = "web dump: page " + std::to_string(p0 + i + 1) + " has " + std::to_string(c) +
                            " out-links, more than the maximum " + std::to_string(m.max_out);
                return -1;
            }
            row_ptr[p0 + i + 1] = c;
        }
        return 0;
    });
    if (rc) {
        unmap_files(dir, m, files);
        for (auto& e : errors) if (!e.empty()) return fail(e);
        return fail("web dump: read failed");
    }
    row_ptr[0] = 0;
    for (uint64_t p = 0; p < m.pages; ++p) {
        const uint64_t s = (uint64_t)row_ptr[p] + row_ptr[p + 1];
        if (s > 0xffffffffull) { unmap_files(dir, m, files); return fail("web dump: more than 2^32 - 1 links"); }
        row_ptr[p + 1] = (uint32_t)s;
    }
    *edges = row_ptr[m.pages];
    if (col) {
        if (col_capacity < *edges) { unmap_files(dir, m, files); return fail("ua_webdump_read: col capacity too small"); }
        rc = for_files(files, [&](uint64_t f) {
            const uint32_t* r;
            uint32_t n;
            if (map_file(dir, m, f, &r, &n, &errors[f])) return -1;
            const uint64_t p0 = f * m.per_file;
            for (uint32_t i = 0; i < n; ++i) {
                const uint32_t* rec = r + i * stride;
                uint32_t* dst = col + row_ptr[p0 + i];
                for (uint32_t k = 0; k < rec[0]; ++k) {
                    const uint32_t t = rec[1 + k];
                    if (t == 0 || t > m.pages) {
                        errors[f] = "web dump: invalid out-link " + std::to_string(t) + " on page " +
                                    std::to_string(p0 + i + 1);
                        return -1;
                    }
                    dst[k] = t - 1;
                }
            }
            return 0;
        });
        if (rc) {
            unmap_files(dir, m, files);
            for (auto& e : errors) if (!e.empty()) return fail(e);
            return fail("web dump: read failed");
        }
    }
    unmap_files(dir, m, files);
    return 0;
}

// Writes an out-link CSR (0-based targets) in the reference's format; max_out must bound every
// page's out-degree.
extern "C" int ua_webdump_write(const char* dir_, uint32_t pages, uint32_t max_out, uint32_t per_file,
                                const uint32_t* row_ptr, const uint32_t* col) {
    if (!dir_ || !row_ptr || !pages || !per_file) return fail("ua_webdump_write: invalid arguments");
    const std::string dir = dir_;
    if (mkdir(dir.c_str(), 0755) != 0 && errno != EEXIST) return fail("ua_webdump_write: cannot create " + dir);
    if (mkdir((dir + "/web").c_str(), 0755) != 0 && errno != EEXIST)
        return fail("ua_webdump_write: cannot create " + dir + "/web");
    {
        FILE* fp = fopen((dir + "/web_dump_metadata").c_str(), "wb");
        if (!fp) return fail("ua_webdump_write: cannot write the metadata");
        const uint32_t v[3] = {pages, max_out, per_file};
        const bool ok = fwrite(v, 4, 3, fp) == 3;
        if (fclose(fp) != 0 || !ok) return fail("ua_webdump_write: cannot write the metadata");
    }
    const uint64_t files = (pages + (uint64_t)per_file - 1) / per_file;
    std::vector<std::string> errors(files);
    const int rc = for_files(files, [&](uint64_t f) {
        const uint64_t p0 = f * per_file;
        const uint64_t n = std::min<uint64_t>(per_file, pages - p0);
        std::vector<uint32_t> buf(n * ((uint64_t)max_out + 1), 0u);
        for (uint64_t i = 0; i < n; ++i) {
            const uint32_t b = row_ptr[p0 + i], e = row_ptr[p0 + i + 1];
            if (e - b > max_out) { errors[f] = "ua_webdump_write: a page has more out-links than max_out"; return -1; }
            uint32_t* rec = buf.data() + i * ((uint64_t)max_out + 1);
            rec[0] = e - b;
            for (uint32_t k = b; k < e; ++k) rec[1 + k - b] = col[k] + 1;
        }
        const std::string path = page_file(dir, p0);
        FILE* fp = fopen(path.c_str(), "wb");
        if (!fp) { errors[f] = "ua_webdump_write: cannot write " + path; return -1; }
        const bool ok = fwrite(buf.data(), 4, buf.size(), fp) == buf.size();
        if (fclose(fp) != 0 || !ok) { errors[f] = "ua_webdump_write: cannot write " + path; return -1; }
        return 0;
    });
    if (rc) {
        for (auto& e : errors) if (!e.empty()) return fail(e);
        return fail("ua_webdump_write failed");
    }
    return 0;
}

// A pageRank instance over a web dump (the graph address space is built once from it).
extern "C" void* ua_pr_create_from_webdump(const char* dir, uint32_t pages_per_subtask, float damping) {
    uint32_t pages, max_out, per_file;
    if (ua_webdump_meta(dir, &pages, &max_out, &per_file)) return nullptr;
    std::vector<uint32_t> rp((size_t)pages + 1);
    uint64_t edges = 0;
    if (ua_webdump_read(dir, rp.data(), nullptr, 0, &edges)) return nullptr;
    std::vector<uint32_t> col(std::max<uint64_t>(edges, 1));
    if (ua_webdump_read(dir, rp.data(), col.data(), col.size(), &edges)) return nullptr;
    return ua_pr_create(pages, rp.data(), col.data(), pages_per_subtask, damping);
}
