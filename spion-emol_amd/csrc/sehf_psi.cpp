// sehf_psi.cpp -- per-energy psi files (layout in sehf_psi.h).
#include "sehf_psi.h"

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>

namespace sehf {

namespace {

constexpr char kMagic[2][8] = {{'P', 'S', 'I', '_', 'A', 'S', 'Y', 'M'}, {'P', 'S', 'I', '_', 'M', 'A', 'I', 'N'}};
constexpr const char* kName[2] = {"psi_asym.bin", "psi_main.bin"};
constexpr int32_t kVersion = 1;
constexpr int64_t kHeader = 32;

[[noreturn]] void fail(const std::string& what, const std::string& path) {
    throw std::runtime_error("psi store: " + what + " " + path + (errno ? std::string(": ") + std::strerror(errno) : ""));
}

void mkdirs(const std::string& path) {
    std::string cur;
    for (size_t i = 0; i <= path.size(); ++i) {
        if (i == path.size() || path[i] == '/') {
            if (!cur.empty() && ::mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) fail("cannot create", cur);
            errno = 0;
        }
        if (i < path.size()) cur += path[i];
    }
}

void write_all(int fd, const void* p, size_t bytes, int64_t off, const std::string& path) {
    const char* c = static_cast<const char*>(p);
    while (bytes > 0) {
        const ssize_t w = ::pwrite(fd, c, bytes, static_cast<off_t>(off));
        if (w < 0) {
            if (errno == EINTR) continue;
            fail("write failed on", path);
        }
        c += w; bytes -= static_cast<size_t>(w); off += w;
    }
}

void read_all(int fd, void* p, size_t bytes, int64_t off, const std::string& path) {
    char* c = static_cast<char*>(p);
    while (bytes > 0) {
        const ssize_t r = ::pread(fd, c, bytes, static_cast<off_t>(off));
        if (r < 0) {
            if (errno == EINTR) continue;
            fail("read failed on", path);
        }
        if (r == 0) { errno = 0; fail("file ends early:", path); }
        c += r; bytes -= static_cast<size_t>(r); off += r;
    }
}

struct Header {
    char magic[8];
    int32_t version, n, first, n_psi;
    int64_t reserved;
};
static_assert(sizeof(Header) == kHeader, "psi file header is 32 bytes");

std::string marker_text(double k, const PsiLayout& l, bool tail, bool main) {
    char buf[256];
    std::snprintf(buf, sizeof buf, "k %.17g\nNr %d\nN_psi %d\nn_asym %d\nasym_offset %d\nn_main %d\n", k, l.Nr, l.Np,
                  tail ? l.n_tail : 0, tail ? l.tail_lo : 0, main ? l.n_main : 0);
    return buf;
}

}  // namespace

std::string psi_dir(const std::string& root, int ik) {
    char buf[32];
    std::snprintf(buf, sizeof buf, "/psi_ik%04d", ik);
    return root + buf;
}

PsiWriter::PsiWriter(const std::string& root, int ik0, int nb, const double* ks, const PsiLayout& lay, bool tail,
                     bool main)
    : lay_(lay) {
    on_[0] = tail && lay.n_tail > 0;
    on_[1] = main && lay.n_main > 0;
    e_.resize(nb);
    try {
        open_all(root, ik0, ks);
    } catch (...) {
        cleanup();   // a throwing constructor runs no destructor
        throw;
    }
}

void PsiWriter::open_all(const std::string& root, int ik0, const double* ks) {
    const PsiLayout& lay = lay_;
    const int nb = static_cast<int>(e_.size());
    const int64_t mat = static_cast<int64_t>(lay.Np) * lay.Np * 8;
    for (int e = 0; e < nb; ++e) {
        One& o = e_[e];
        o.dir = psi_dir(root, ik0 + e);
        o.k = ks[e];
        errno = 0;
        mkdirs(o.dir);
        // an older marker goes first: from here on the directory is incomplete
        if (::unlink((o.dir + "/__SUCCESS__").c_str()) != 0 && errno != ENOENT) fail("cannot remove", o.dir + "/__SUCCESS__");
        for (int f = 0; f < 2; ++f) {
            if (!on_[f]) continue;
            o.fin[f] = o.dir + "/" + kName[f];
            o.tmp[f] = o.fin[f] + ".tmp" + std::to_string(::getpid());
            errno = 0;
            o.fd[f] = ::open(o.tmp[f].c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
            if (o.fd[f] < 0) fail("cannot create", o.tmp[f]);
            Header h{};
            std::memcpy(h.magic, kMagic[f], 8);
            h.version = kVersion;
            h.n = f == 0 ? lay.n_tail : lay.n_main;
            h.first = f == 0 ? lay.tail_lo : 0;
            h.n_psi = lay.Np;
            h.reserved = 0;
            write_all(o.fd[f], &h, sizeof h, 0, o.tmp[f]);
            if (::ftruncate(o.fd[f], static_cast<off_t>(kHeader + h.n * mat)) != 0) fail("cannot size", o.tmp[f]);
        }
    }
}

PsiWriter::~PsiWriter() { cleanup(); }

void PsiWriter::cleanup() {
    for (One& o : e_)
        for (int f = 0; f < 2; ++f) {
            if (o.fd[f] >= 0) ::close(o.fd[f]);
            o.fd[f] = -1;
            if (!done_ && !o.tmp[f].empty()) ::unlink(o.tmp[f].c_str());
        }
}

bool PsiWriter::wants(int file, int n) const {
    if (!on_[file]) return false;
    return file == 0 ? (n >= lay_.tail_lo && n < lay_.tail_lo + lay_.n_tail) : (n >= 0 && n < lay_.n_main);
}

void PsiWriter::put(int file, int n, const double* batch) {
    if (!wants(file, n)) return;
    const size_t mat = static_cast<size_t>(lay_.Np) * lay_.Np;
    const int64_t off = kHeader + static_cast<int64_t>(n - (file == 0 ? lay_.tail_lo : 0)) * mat * 8;
    for (size_t e = 0; e < e_.size(); ++e) write_all(e_[e].fd[file], batch + e * mat, mat * 8, off, e_[e].tmp[file]);
}

void PsiWriter::commit() {
    for (One& o : e_) {
        for (int f = 0; f < 2; ++f) {
            if (o.fd[f] < 0) continue;
            errno = 0;
            if (::fsync(o.fd[f]) != 0) fail("fsync failed on", o.tmp[f]);
            ::close(o.fd[f]);
            o.fd[f] = -1;
            if (::rename(o.tmp[f].c_str(), o.fin[f].c_str()) != 0) fail("cannot rename", o.tmp[f]);
        }
        const std::string txt = marker_text(o.k, lay_, on_[0], on_[1]);
        const std::string mk = o.dir + "/__SUCCESS__";
        errno = 0;
        const int fd = ::open(mk.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
        if (fd < 0) fail("cannot create", mk);
        write_all(fd, txt.data(), txt.size(), 0, mk);
        ::close(fd);
    }
    done_ = true;
}

PsiReader::PsiReader(const std::string& root, int ik0, int nb, const double* ks, const PsiLayout& lay, bool need_main)
    : lay_(lay), main_(need_main) {
    e_.resize(nb);
    try {
        open_all(root, ik0, ks);
    } catch (...) {
        close_all();   // a throwing constructor runs no destructor
        throw;
    }
}

void PsiReader::open_all(const std::string& root, int ik0, const double* ks) {
    const PsiLayout& lay = lay_;
    const bool need_main = main_;
    const int nb = static_cast<int>(e_.size());
    for (int e = 0; e < nb; ++e) {
        const std::string dir = psi_dir(root, ik0 + e);
        errno = 0;
        FILE* m = std::fopen((dir + "/__SUCCESS__").c_str(), "r");
        if (!m) fail("no completed psi files (__SUCCESS__ missing) in", dir);
        char key[32];
        double kv = NAN;
        while (std::fscanf(m, "%31s", key) == 1) {
            if (std::strcmp(key, "k") == 0) { if (std::fscanf(m, "%lf", &kv) != 1) break; }
            else { int skip; if (std::fscanf(m, "%d", &skip) != 1) break; }
        }
        std::fclose(m);
        errno = 0;
        if (!(kv == ks[e]))
            fail("energy mismatch (k " + std::to_string(kv) + " on file, " + std::to_string(ks[e]) + " asked) in", dir);
        One& o = e_[e];
        for (int f = 0; f < 2; ++f) {
            if (f == 1 && !need_main) continue;
            o.path[f] = dir + "/" + kName[f];
            errno = 0;
            o.fd[f] = ::open(o.path[f].c_str(), O_RDONLY);
            if (o.fd[f] < 0) fail("cannot open", o.path[f]);
            Header h{};
            read_all(o.fd[f], &h, sizeof h, 0, o.path[f]);
            errno = 0;
            if (std::memcmp(h.magic, kMagic[f], 8) != 0 || h.version != kVersion) fail("not a psi file of this layout:", o.path[f]);
            if (h.n_psi != lay.Np) fail("N_psi " + std::to_string(h.n_psi) + " does not match the problem in", o.path[f]);
            if (f == 0 && (h.first != lay.tail_lo || h.n != lay.n_tail))
                fail("tail range [" + std::to_string(h.first) + ", +" + std::to_string(h.n) + ") does not match the grid in",
                     o.path[f]);
            if (f == 1 && (h.first != 0 || h.n < lay.n_main))
                fail("main range holds " + std::to_string(h.n) + " points, the dipoles need " + std::to_string(lay.n_main) +
                         ", in",
                     o.path[f]);
            struct stat sb;
            const int64_t need = kHeader + static_cast<int64_t>(h.n) * lay.Np * lay.Np * 8;
            if (::fstat(o.fd[f], &sb) != 0 || sb.st_size < need) fail("file is shorter than its header says:", o.path[f]);
        }
    }
}

PsiReader::~PsiReader() { close_all(); }

void PsiReader::close_all() {
    for (One& o : e_)
        for (int f = 0; f < 2; ++f) {
            if (o.fd[f] >= 0) ::close(o.fd[f]);
            o.fd[f] = -1;
        }
}

void PsiReader::get(int n, double* batch) const {
    const size_t mat = static_cast<size_t>(lay_.Np) * lay_.Np;
    int f = -1;
    int64_t idx = 0;
    if (main_ && n >= 0 && n < lay_.n_main) { f = 1; idx = n; }
    else if (n >= lay_.tail_lo && n < lay_.tail_lo + lay_.n_tail) { f = 0; idx = n - lay_.tail_lo; }
    else throw std::runtime_error("psi store: point " + std::to_string(n) + " is in neither stored range");
    for (size_t e = 0; e < e_.size(); ++e)
        read_all(e_[e].fd[f], batch + e * mat, mat * 8, kHeader + idx * static_cast<int64_t>(mat) * 8, e_[e].path[f]);
}

}  // namespace sehf
