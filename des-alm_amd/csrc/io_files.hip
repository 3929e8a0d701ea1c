// io_files.hip -- host-side readers and writers of the reference's files
// (no device code):
//   grid              src/init.F90:2226-2241 (x, y, z and the wall distance)
//   solu              src/init.F90:2384-2460 (q, xnut, qn)
//   solufiles/solu.N  src/main.F90:343-372 (restart output, per rank)
//   filestat          src/init.F90:266-277, src/main.F90:379-386
//   ind3dmg.dat       src/init.F90:176-228 (the input deck)
//
// grid and solu are Fortran unformatted sequential files: every record is
// framed by a 4-byte little-endian byte count before and after the payload.
// A record longer than 2^31 - 9 bytes is split into subrecords; the leading
// count of a subrecord is negative when the record continues in the next one
// and the trailing count is negative when the subrecord continues a previous
// one (the gfortran convention).  The reference computes and writes REAL*4
// (src/precision.F90:11); an FP64 build writes REAL*8.  Both are accepted.
//
// The deck and filestat are list-directed text.  ListReader implements the
// part of list-directed input those files use: one READ statement starts on
// a new line and may continue over following lines; values are separated by
// blanks, commas or line ends; "r*v" repeats v r times, "r*" and an empty
// field between commas are null values (the item keeps its value); a slash
// ends the statement; the rest of the last line read is skipped.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/desalm.h"
#include "host_error.h"

namespace desalm {
namespace {

constexpr size_t kMaxSubrecord = 2147483639u;  // 2^31 - 9 bytes

struct File {
    FILE* f = nullptr;
    explicit File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
    File(const File&) = delete;
    File& operator=(const File&) = delete;
};

// One logical record (all of its subrecords) into buf.
bool read_record(FILE* f, std::vector<unsigned char>& buf, std::string& err) {
    buf.clear();
    for (int sub = 0;; ++sub) {
        int32_t head = 0, tail = 0;
        if (std::fread(&head, 4, 1, f) != 1) {
            err = sub ? "truncated record" : "unexpected end of file";
            return false;
        }
        const uint32_t len = head < 0 ? (uint32_t)(-(int64_t)head) : (uint32_t)head;
        // refuse a length the rest of the file cannot hold before allocating it
        const long here = std::ftell(f);
        if (here < 0 || std::fseek(f, 0, SEEK_END) != 0) {
            err = "cannot seek in the file";
            return false;
        }
        const long end = std::ftell(f);
        if (std::fseek(f, here, SEEK_SET) != 0 || end < 0) {
            err = "cannot seek in the file";
            return false;
        }
        if ((int64_t)len + 4 > (int64_t)end - (int64_t)here) {
            err = "truncated record";
            return false;
        }
        const size_t at = buf.size();
        buf.resize(at + len);
        if (len && std::fread(buf.data() + at, 1, len, f) != len) {
            err = "truncated record";
            return false;
        }
        if (std::fread(&tail, 4, 1, f) != 1) {
            err = "truncated record";
            return false;
        }
        const uint32_t tlen = tail < 0 ? (uint32_t)(-(int64_t)tail) : (uint32_t)tail;
        if (tlen != len) {
            err = "record length markers disagree (not a Fortran unformatted sequential file?)";
            return false;
        }
        if ((tail < 0) != (sub > 0)) {
            err = "subrecord markers out of order";
            return false;
        }
        if (head >= 0) return true;
    }
}

bool write_record(FILE* f, const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    size_t done = 0;
    bool first = true;
    do {
        const size_t len = std::min(n - done, kMaxSubrecord);
        const bool more = done + len < n;
        const int32_t head = more ? -(int32_t)len : (int32_t)len;
        const int32_t tail = first ? (int32_t)len : -(int32_t)len;
        if (std::fwrite(&head, 4, 1, f) != 1) return false;
        if (len && std::fwrite(b + done, 1, len, f) != len) return false;
        if (std::fwrite(&tail, 4, 1, f) != 1) return false;
        done += len;
        first = false;
    } while (done < n);
    return true;
}

// Reads one record of `count` reals into out[stride * l] (l < count).
// real_bytes: 4, 8, or 0 (taken from this record's length, then kept).
bool read_reals(FILE* f, std::vector<unsigned char>& buf, long count, int& real_bytes, double* out,
                long stride, const char* what, std::string& err) {
    if (!read_record(f, buf, err)) {
        err = std::string(what) + ": " + err;
        return false;
    }
    if (real_bytes == 0 && count > 0 && (buf.size() == 4u * (size_t)count || buf.size() == 8u * (size_t)count))
        real_bytes = (int)(buf.size() / (size_t)count);
    // a record whose size fits neither REAL*4 nor REAL*8 (an empty one
    // included) leaves real_bytes undetected: never accept it as 0 * count
    if ((real_bytes != 4 && real_bytes != 8) || buf.size() != (size_t)real_bytes * (size_t)count) {
        err = std::string(what) + ": record holds " + std::to_string(buf.size()) + " bytes, expected " +
              std::to_string(count) + " reals";
        return false;
    }
    if (real_bytes == 4) {
        for (long l = 0; l < count; l++) {
            float v;
            std::memcpy(&v, buf.data() + 4 * l, 4);
            out[stride * l] = (double)v;
        }
    } else {
        for (long l = 0; l < count; l++) std::memcpy(&out[stride * l], buf.data() + 8 * l, 8);
    }
    return true;
}

// Writes count reals in[stride * l] as one record of REAL*4 or REAL*8.
bool write_reals(FILE* f, std::vector<unsigned char>& buf, long count, int real_bytes, const double* in,
                 long stride) {
    buf.resize((size_t)real_bytes * (size_t)count);
    if (real_bytes == 4) {
        for (long l = 0; l < count; l++) {
            const float v = (float)in[stride * l];
            std::memcpy(buf.data() + 4 * l, &v, 4);
        }
    } else {
        for (long l = 0; l < count; l++) std::memcpy(buf.data() + 8 * l, &in[stride * l], 8);
    }
    return write_record(f, buf.data(), buf.size());
}

bool zone_sizes(int32_t nzone, const int32_t* dims, std::vector<long>& n, std::string& err) {
    if (nzone < 1 || !dims) {
        err = "nzone must be >= 1 with dims[3 * nzone]";
        return false;
    }
    n.resize(nzone);
    for (int z = 0; z < nzone; z++) {
        if (dims[3 * z] < 1 || dims[3 * z + 1] < 1 || dims[3 * z + 2] < 1) {
            err = "zone " + std::to_string(z + 1) + ": node counts must be positive";
            return false;
        }
        n[z] = (long)dims[3 * z] * dims[3 * z + 1] * dims[3 * z + 2];
    }
    return true;
}

bool real_bytes_ok(int rb) { return rb == 0 || rb == 4 || rb == 8; }

// ---- list-directed input ----
struct Item {
    bool null = true;
    std::string text;
};

class ListReader {
  public:
    bool open(const char* path, std::string& err) {
        File in(path, "r");
        if (!in.f) {
            err = std::string("cannot open ") + path;
            return false;
        }
        std::string cur;
        int ch;
        while ((ch = std::fgetc(in.f)) != EOF) {
            if (ch == '\n') {
                lines_.push_back(cur);
                cur.clear();
            } else if (ch != '\r') {
                cur.push_back((char)ch);
            }
        }
        if (!cur.empty()) lines_.push_back(cur);
        return true;
    }

    // One READ statement of n items.
    bool statement(int n, std::vector<Item>& items, std::string& err) {
        items.assign(n, Item());
        int got = 0;
        // true at the start of the statement and after a comma: a comma there is a null value
        bool comma_is_null = true;
        while (got < n) {
            if (line_ >= lines_.size()) {
                err = "end of file in the middle of a list-directed read";
                return false;
            }
            const std::string& s = lines_[line_];
            while (pos_ < s.size() && got < n) {
                const char c = s[pos_];
                if (c == ' ' || c == '\t') {
                    pos_++;
                } else if (c == ',') {
                    if (comma_is_null) items[got++].null = true;
                    comma_is_null = true;
                    pos_++;
                } else if (c == '/') {
                    finish();
                    return true;  // the remaining items keep their values
                } else {
                    size_t e = pos_;
                    while (e < s.size() && s[e] != ' ' && s[e] != '\t' && s[e] != ',' && s[e] != '/') e++;
                    const std::string tok = s.substr(pos_, e - pos_);
                    pos_ = e;
                    comma_is_null = false;
                    const size_t star = tok.find('*');
                    if (star != std::string::npos && star > 0 && tok.find_first_not_of("0123456789") == star) {
                        const long r = std::atol(tok.substr(0, star).c_str());
                        const std::string v = tok.substr(star + 1);
                        if (r < 1) {
                            err = "bad repeat count in '" + tok + "'";
                            return false;
                        }
                        for (long t = 0; t < r && got < n; t++) {
                            items[got].null = v.empty();
                            items[got].text = v;
                            got++;
                        }
                    } else {
                        items[got].null = false;
                        items[got].text = tok;
                        got++;
                    }
                }
            }
            if (got < n) {
                line_++;
                pos_ = 0;
            }
        }
        finish();
        return true;
    }

  private:
    void finish() {
        line_++;
        pos_ = 0;
    }
    std::vector<std::string> lines_;
    size_t line_ = 0, pos_ = 0;
};

}  // namespace
}  // namespace desalm

namespace desalm {
namespace {

bool to_int(const Item& it, int32_t& v) {
    if (it.null) return true;
    const char* s = it.text.c_str();
    char* end = nullptr;
    const long r = std::strtol(s, &end, 10);
    if (end == s || *end != '\0' || r < INT32_MIN || r > INT32_MAX) return false;
    v = (int32_t)r;
    return true;
}

// Fortran real forms: 1, 1., .5, 1.0E-08, 1.0D-08, 1.0d0
bool to_real(const Item& it, double& v) {
    if (it.null) return true;
    std::string t = it.text;
    for (char& c : t)
        if (c == 'd' || c == 'D' || c == 'q' || c == 'Q') c = 'E';
    const char* s = t.c_str();
    char* end = nullptr;
    const double r = std::strtod(s, &end);
    if (end == s || *end != '\0') return false;
    v = r;
    return true;
}

// .true., T, .T., true, .false., F ... (the first letter after an optional period)
bool to_logical(const Item& it, int32_t& v) {
    if (it.null) return true;
    const std::string& t = it.text;
    const size_t p = (!t.empty() && t[0] == '.') ? 1 : 0;
    if (p >= t.size()) return false;
    const char c = t[p];
    if (c == 'T' || c == 't') v = 1;
    else if (c == 'F' || c == 'f') v = 0;
    else return false;
    return true;
}

struct Deck {
    ListReader rd;
    std::vector<Item> it;
    std::string err;
    int stmt = 0;
    bool next(int n, const char* what) {
        stmt++;
        if (!rd.statement(n, it, err)) {
            err = std::string("read ") + std::to_string(stmt) + " (" + what + "): " + err;
            return false;
        }
        return true;
    }
    bool bad(int k, const char* what) {
        err = std::string("read ") + std::to_string(stmt) + " (" + what + "): cannot convert '" + it[k].text + "'";
        return false;
    }
    bool i(int k, int32_t& v, const char* what) { return to_int(it[k], v) || bad(k, what); }
    bool r(int k, double& v, const char* what) { return to_real(it[k], v) || bad(k, what); }
    bool l(int k, int32_t& v, const char* what) { return to_logical(it[k], v) || bad(k, what); }
};

bool parse_deck(Deck& D, int32_t ng, desalm_deck* d) {
    desalm_params& p = d->params;
    int32_t itm[DESALM_MAX_LEVELS] = {};
    if (!D.next(2, "ns, icycle") || !D.i(0, d->ns, "ns") || !D.i(1, d->icycle, "icycle")) return false;
    if (!D.next(3, "img, jmg, kmg") || !D.i(0, p.img, "img") || !D.i(1, p.jmg, "jmg") || !D.i(2, p.kmg, "kmg"))
        return false;
    if (!D.next(3, "icrs, jcrs, kcrs") || !D.i(0, p.icrs, "icrs") || !D.i(1, p.jcrs, "jcrs") ||
        !D.i(2, p.kcrs, "kcrs"))
        return false;
    if (!D.next(5, "nt2, delti, it_min, er_min, eo_min") || !D.i(0, d->nt2, "nt2") || !D.r(1, p.delti, "delti") ||
        !D.i(2, p.it_min, "it_min") || !D.r(3, p.er_min, "er_min") || !D.r(4, p.eo_min, "eo_min"))
        return false;
    if (!D.next(ng, "iter(1:ng)")) return false;
    for (int n = 0; n < ng; n++)
        if (!D.i(n, p.iter[n], "iter")) return false;
    if (!D.next(ng, "itm(1:ng)")) return false;
    for (int n = 0; n < ng; n++)
        if (!D.i(n, itm[n], "itm")) return false;
    p.itm = itm[0];
    for (int n = 0; n < ng; n++) d->itm[n] = itm[n];
    if (!D.next(4, "ren, beta, icn, icnw") || !D.r(0, p.ren, "ren") || !D.r(1, p.beta, "beta") ||
        !D.i(2, d->icn, "icn") || !D.i(3, d->icnw, "icnw"))
        return false;
    if (!D.next(2, "cfl1, vnn1") || !D.r(0, p.cfl1, "cfl1") || !D.r(1, p.vnn1, "vnn1")) return false;
    // irk, alfa(1:irk): the count is the first item of the same statement
    {
        ListReader save = D.rd;
        if (!D.next(1, "irk") || !D.i(0, d->irk, "irk")) return false;
        if (d->irk < 0 || d->irk > 4) {
            D.err = "irk = " + std::to_string(d->irk) + " outside 0..4 (alfa(4), src/global_param.F90:28)";
            return false;
        }
        D.rd = save;
        D.stmt--;
        if (!D.next(1 + d->irk, "irk, alfa(1:irk)")) return false;
        for (int k = 0; k < d->irk; k++)
            if (!D.r(1 + k, d->alfa[k], "alfa")) return false;
    }
    if (!D.next(4, "eps(1:4)")) return false;
    for (int k = 0; k < 4; k++)
        if (!D.r(k, d->eps[k], "eps")) return false;
    if (!D.next(4, "itk, cdes, ckin, cein") || !D.i(0, p.itk, "itk") || !D.r(1, d->cdes, "cdes") ||
        !D.r(2, d->ckin, "ckin") || !D.r(3, d->cein, "cein"))
        return false;
    if (!D.next(2, "cfl2, vnn2") || !D.r(0, p.cfl2, "cfl2") || !D.r(1, p.vnn2, "vnn2")) return false;
    if (!D.next(6, "ep(1:2, 1:3)")) return false;
    for (int j = 0; j < 2; j++)
        for (int k = 0; k < 3; k++)
            if (!D.r(3 * j + k, p.ep[j][k], "ep")) return false;
    if (!D.next(1, "checkpoint") || !D.i(0, d->checkpoint, "checkpoint")) return false;
    if (!D.next(1, "pdiss_coef") || !D.r(0, p.pdiss_coef, "pdiss_coef")) return false;
    if (!D.next(1, "monitor_num_points") || !D.i(0, d->monitor_num_points, "monitor_num_points")) return false;
    if (d->monitor_num_points < 0 || d->monitor_num_points > 10) {
        D.err = "monitor_num_points = " + std::to_string(d->monitor_num_points) +
                " outside 0..10 (monitor_point_ijk(10, 4), src/global_param.F90:68)";
        return false;
    }
    if (!D.next(4 * d->monitor_num_points, "monitor_point_ijk")) return false;
    for (int j = 0; j < d->monitor_num_points; j++)
        for (int k = 0; k < 4; k++)
            if (!D.i(4 * j + k, d->monitor_point_ijk[j][k], "monitor_point_ijk")) return false;
    if (!D.next(6, "btype(1:6)")) return false;
    for (int f = 0; f < 6; f++)
        if (!D.i(f, p.btype[f], "btype")) return false;
    if (!D.next(6, "bdir(1:6)")) return false;
    for (int f = 0; f < 6; f++)
        if (!D.i(f, p.bdir[f], "bdir")) return false;
    if (!D.next(4, "ifix, jfix, kfix, nzfix") || !D.i(0, p.ifix, "ifix") || !D.i(1, p.jfix, "jfix") ||
        !D.i(2, p.kfix, "kfix") || !D.i(3, d->nzfix, "nzfix"))
        return false;
    if (!D.next(1, "length_scale") || !D.r(0, d->length_scale, "length_scale")) return false;
    if (!D.next(1, "vel_infty") || !D.r(0, d->vel_infty, "vel_infty")) return false;
    if (!D.next(1, "turbines") || !D.l(0, d->turbines, "turbines")) return false;
    if (!D.next(1, "node_indexes_load") || !D.l(0, d->node_indexes_load, "node_indexes_load")) return false;
    if (!D.next(1, "n_tur") || !D.i(0, d->n_tur, "n_tur")) return false;
    if (d->n_tur < 0 || d->n_tur > DESALM_DECK_MAX_TURBINES) {
        D.err = "n_tur = " + std::to_string(d->n_tur) + " outside 0.." + std::to_string(DESALM_DECK_MAX_TURBINES);
        return false;
    }
    for (int t = 0; t < d->n_tur; t++) {
        desalm_deck_turbine& T = d->turbine[t];
        if (!D.next(7, "itur1, itur2, jtur1, jtur2, ktur1, ktur2, tur_rad") || !D.i(0, T.itur1, "itur1") ||
            !D.i(1, T.itur2, "itur2") || !D.i(2, T.jtur1, "jtur1") || !D.i(3, T.jtur2, "jtur2") ||
            !D.i(4, T.ktur1, "ktur1") || !D.i(5, T.ktur2, "ktur2") || !D.r(6, T.tur_rad, "tur_rad"))
            return false;
        if (!D.next(2, "hub_rad, TSR") || !D.r(0, T.hub_rad, "hub_rad") || !D.r(1, T.tsr, "TSR")) return false;
        if (!D.next(2, "nblades, nelems") || !D.i(0, T.nblades, "nblades") || !D.i(1, T.nelems, "nelems"))
            return false;
    }
    // blanking of the single zone (nzone = 1, nba_max = 1: src/global.F90:20-21)
    if (!D.next(1, "nzblanking") || !D.i(0, d->nzblanking, "nzblanking")) return false;
    if (d->nzblanking < 0 || d->nzblanking > 1) {
        D.err = "nzblanking = " + std::to_string(d->nzblanking) + " outside 0..1 (nba_max = 1, src/global.F90:21)";
        return false;
    }
    if (d->nzblanking == 1) {
        if (!D.next(6, "blktype(1:6)")) return false;
        for (int f = 0; f < 6; f++)
            if (!D.i(f, d->blktype[f], "blktype")) return false;
        if (!D.next(6, "blanking(1:3, 1:2)")) return false;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 2; b++)
                if (!D.i(2 * a + b, d->blanking[a][b], "blanking")) return false;
    }
    return true;
}

bool write_text(const char* path, const std::string& s) {
    File out(path, "w");
    if (!out.f) return false;
    return std::fwrite(s.data(), 1, s.size(), out.f) == s.size();
}

}  // namespace
}  // namespace desalm

using namespace desalm;

extern "C" {

int desalm_read_grid_file(const char* path, int32_t nzone, const int32_t* dims, int32_t turbulence,
                          int32_t real_bytes, double* x, double* y, double* z, double* wd) {
    std::string err;
    std::vector<long> n;
    if (!path || !x || !y || !z || (turbulence && !wd)) return host_fail("desalm_read_grid_file: null argument");
    if (!real_bytes_ok(real_bytes)) return host_fail("desalm_read_grid_file: real_bytes must be 0, 4 or 8");
    if (!zone_sizes(nzone, dims, n, err)) return host_fail("desalm_read_grid_file: " + err);
    File in(path, "rb");
    if (!in.f) return host_fail(std::string("desalm_read_grid_file: cannot open ") + path);
    std::vector<unsigned char> buf;
    int rb = real_bytes;
    long off = 0;
    for (int zn = 0; zn < nzone; zn++) {
        const std::string at = std::string(path) + ", zone " + std::to_string(zn + 1);
        if (!read_reals(in.f, buf, n[zn], rb, x + off, 1, "x", err) ||
            !read_reals(in.f, buf, n[zn], rb, y + off, 1, "y", err) ||
            !read_reals(in.f, buf, n[zn], rb, z + off, 1, "z", err) ||
            (turbulence && !read_reals(in.f, buf, n[zn], rb, wd + off, 1, "wall distance", err)))
            return host_fail("desalm_read_grid_file: " + at + ": " + err);
        off += n[zn];
    }
    return 0;
}

int desalm_read_solu_file(const char* path, int32_t nzone, const int32_t* dims, int32_t turbulence,
                          int32_t real_bytes, double* q, double* qn, double* xnut) {
    std::string err;
    std::vector<long> n;
    if (!path || !q || !qn || (turbulence && !xnut)) return host_fail("desalm_read_solu_file: null argument");
    if (!real_bytes_ok(real_bytes)) return host_fail("desalm_read_solu_file: real_bytes must be 0, 4 or 8");
    if (!zone_sizes(nzone, dims, n, err)) return host_fail("desalm_read_solu_file: " + err);
    File in(path, "rb");
    if (!in.f) return host_fail(std::string("desalm_read_solu_file: cannot open ") + path);
    std::vector<unsigned char> buf;
    int rb = real_bytes;
    long off = 0;
    const int me = 5;
    for (int zn = 0; zn < nzone; zn++) {
        const std::string at = std::string(path) + ", zone " + std::to_string(zn + 1);
        const long nz = n[zn];
        if (turbulence) {
            // me records of q, one of xnut, me of qn (src/init.F90:2404-2410)
            for (int m = 0; m < me; m++)
                if (!read_reals(in.f, buf, nz, rb, q + me * off + m, me, "q", err))
                    return host_fail("desalm_read_solu_file: " + at + ": " + err);
            if (!read_reals(in.f, buf, nz, rb, xnut + off, 1, "xnut", err))
                return host_fail("desalm_read_solu_file: " + at + ": " + err);
            for (int m = 0; m < me; m++)
                if (!read_reals(in.f, buf, nz, rb, qn + me * off + m, me, "qn", err))
                    return host_fail("desalm_read_solu_file: " + at + ": " + err);
        } else {
            // one record of q(1:me) (component after component) and one of qn
            // (src/init.F90:2451-2452); a file written by the restart output
            // (one record per component, src/main.F90:356-372) is read too.
            for (int which = 0; which < 2; which++) {
                double* dst = (which ? qn : q) + me * off;
                const char* nm = which ? "qn" : "q";
                if (!read_record(in.f, buf, err)) return host_fail("desalm_read_solu_file: " + at + ": " + nm + ": " + err);
                int rb1 = rb;
                const size_t all4 = 4u * (size_t)me * (size_t)nz, all8 = 2 * all4;
                const size_t one4 = 4u * (size_t)nz, one8 = 2 * one4;
                bool whole;
                if (rb1 == 0) {
                    if (buf.size() == all4 || buf.size() == all8) rb1 = buf.size() == all4 ? 4 : 8, whole = true;
                    else if (buf.size() == one4 || buf.size() == one8) rb1 = buf.size() == one4 ? 4 : 8, whole = false;
                    else return host_fail("desalm_read_solu_file: " + at + ": " + nm + ": record size does not match the zone");
                } else {
                    whole = buf.size() == (size_t)rb1 * me * (size_t)nz;
                    if (!whole && buf.size() != (size_t)rb1 * (size_t)nz)
                        return host_fail("desalm_read_solu_file: " + at + ": " + nm + ": record size does not match the zone");
                }
                rb = rb1;
                for (int m = 0; m < me; m++) {
                    if (!whole && m > 0 && !read_record(in.f, buf, err))
                        return host_fail("desalm_read_solu_file: " + at + ": " + nm + ": " + err);
                    if (!whole && buf.size() != (size_t)rb * (size_t)nz)
                        return host_fail("desalm_read_solu_file: " + at + ": " + nm + ": record size does not match the zone");
                    const unsigned char* src = buf.data() + (whole ? (size_t)rb * (size_t)nz * m : 0);
                    for (long l = 0; l < nz; l++) {
                        double v;
                        if (rb == 4) {
                            float f;
                            std::memcpy(&f, src + 4 * l, 4);
                            v = f;
                        } else {
                            std::memcpy(&v, src + 8 * l, 8);
                        }
                        dst[me * l + m] = v;
                    }
                }
            }
        }
        off += nz;
    }
    return 0;
}

int desalm_write_solu_file(const char* path, int64_t n1, int32_t turbulence, int32_t real_bytes, const double* q,
                           const double* qn, const double* xnut) {
    if (!path || !q || !qn || (turbulence && !xnut)) return host_fail("desalm_write_solu_file: null argument");
    if (real_bytes != 4 && real_bytes != 8) return host_fail("desalm_write_solu_file: real_bytes must be 4 or 8");
    if (n1 < 1) return host_fail("desalm_write_solu_file: n1 must be positive");
    File out(path, "wb");
    if (!out.f) return host_fail(std::string("desalm_write_solu_file: cannot create ") + path);
    std::vector<unsigned char> buf;
    bool ok = true;
    for (int m = 0; m < 5 && ok; m++) ok = write_reals(out.f, buf, n1, real_bytes, q + m, 5);
    if (ok && turbulence) ok = write_reals(out.f, buf, n1, real_bytes, xnut, 1);
    for (int m = 0; m < 5 && ok; m++) ok = write_reals(out.f, buf, n1, real_bytes, qn + m, 5);
    if (!ok) return host_fail(std::string("desalm_write_solu_file: write failed on ") + path);
    return 0;
}

int desalm_read_filestat(const char* path, int32_t n_tur, double* time0, int32_t* nt1, double* angle) {
    if (!path || !time0 || !nt1 || (n_tur > 0 && !angle)) return host_fail("desalm_read_filestat: null argument");
    ListReader rd;
    std::string err;
    std::vector<Item> it;
    if (!rd.open(path, err)) return host_fail("desalm_read_filestat: " + err);
    if (!rd.statement(2, it, err)) return host_fail(std::string("desalm_read_filestat: time0, nt1: ") + err);
    if (!to_real(it[0], *time0) || !to_int(it[1], *nt1))
        return host_fail("desalm_read_filestat: time0, nt1: cannot convert '" + it[0].text + "', '" + it[1].text + "'");
    const double two_pi = 2.0 * 3.141592653589793;
    for (int t = 0; t < n_tur; t++) {
        if (!rd.statement(1, it, err)) return host_fail("desalm_read_filestat: angle: " + err);
        if (!to_real(it[0], angle[t])) return host_fail("desalm_read_filestat: angle: cannot convert '" + it[0].text + "'");
        if (angle[t] >= two_pi) angle[t] -= two_pi;  // src/init.F90:272-274
    }
    return 0;
}

int desalm_write_filestat(const char* path, double time, int32_t ntime, int32_t n_tur, const double* angle) {
    if (!path || (n_tur > 0 && !angle)) return host_fail("desalm_write_filestat: null argument");
    char line[96];
    std::snprintf(line, sizeof line, " %.17G %d\n", time, ntime);
    std::string s = line;
    for (int t = 0; t < n_tur; t++) {
        std::snprintf(line, sizeof line, " %.17G\n", angle[t]);
        s += line;
    }
    if (!write_text(path, s)) return host_fail(std::string("desalm_write_filestat: cannot write ") + path);
    return 0;
}

int desalm_read_deck(const char* path, int32_t ng, desalm_deck* deck) {
    if (!path || !deck) return host_fail("desalm_read_deck: null argument");
    if (ng < 1 || ng > DESALM_MAX_LEVELS)
        return host_fail("desalm_read_deck: ng must be in 1.." + std::to_string(DESALM_MAX_LEVELS));
    desalm_deck d;
    std::memset(&d, 0, sizeof d);
    d.params.ng = ng;
    d.params.unsteady = 1;  // src/global_param.F90:82
    d.params.des = 1;       // src/global_param.F90:121
    d.params.precision = DESALM_PREC_FP64;
    Deck D;
    if (!D.rd.open(path, D.err)) return host_fail("desalm_read_deck: " + D.err);
    if (!parse_deck(D, ng, &d)) return host_fail(std::string("desalm_read_deck: ") + path + ": " + D.err);
    *deck = d;
    return 0;
}

}  // extern "C"
