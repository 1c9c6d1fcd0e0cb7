// On-disk formats on either side of the feature path, host code:
//   * TIFF input, the subset ImageMatrix::LoadTIFF accepts
//     (src/cmatrix.cpp:85-203): one directory, 8 or 16 bits per sample,
//     1 or 3 samples per pixel, strips;
//   * .sig signature files: signatures::SaveToFile / LoadFromFilep
//     (src/wndchrm_src/signatures.cpp:227-304);
//   * .fit training sets: TrainingSet::SaveToFile / ReadFromFile
//     (src/wndchrm_src/TrainingSet.cpp:339-480);
//   * feature weights: TrainingSet::SaveWeightVector / LoadWeightVector
//     (src/wndchrm_src/TrainingSet.cpp:594-642).
// Writers print with the reference's printf formats, so files are byte-for-
// byte what the reference writes.  Readers follow the reference's parsing.
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <zlib.h>

#include "wndgpu.h"

namespace {

thread_local std::string g_err;

int io_fail(int code, const char *fmt, ...) {
  char buf[768];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

bool read_file(const char *path, std::vector<unsigned char> &out) {
  FILE *f = fopen(path, "rb");
  if (!f) return false;
  unsigned char buf[1 << 16];
  size_t k;
  while ((k = fread(buf, 1, sizeof buf, f)) > 0) out.insert(out.end(), buf, buf + k);
  const bool ok = !ferror(f);
  fclose(f);
  return ok;
}

// ------------------------------------------------------------------ TIFF
struct Tiff {
  const unsigned char *d;
  size_t n;
  bool be;  // "MM"
  bool in(size_t off, size_t len) const { return off <= n && len <= n - off; }
  unsigned u16(size_t o) const { return be ? (d[o] << 8) | d[o + 1] : d[o] | (d[o + 1] << 8); }
  uint32_t u32(size_t o) const {
    return be ? ((uint32_t)d[o] << 24) | ((uint32_t)d[o + 1] << 16) | ((uint32_t)d[o + 2] << 8) | d[o + 3]
              : d[o] | ((uint32_t)d[o + 1] << 8) | ((uint32_t)d[o + 2] << 16) | ((uint32_t)d[o + 3] << 24);
  }
};

struct Field {
  unsigned type = 0;
  uint32_t count = 0;
  size_t off = 0;  // file offset of the values
};

// values of a SHORT or LONG field
bool field_values(const Tiff &t, const Field &f, std::vector<uint32_t> &v) {
  const size_t sz = f.type == 3 ? 2 : f.type == 4 ? 4 : 0;
  if (!sz || !t.in(f.off, (size_t)f.count * sz)) return false;
  v.resize(f.count);
  for (uint32_t i = 0; i < f.count; i++) v[i] = sz == 2 ? t.u16(f.off + 2 * i) : t.u32(f.off + 4 * i);
  return true;
}

// TIFF LZW (compression 5): MSB-first codes of 9..12 bits, 256 = clear,
// 257 = end, and the code width grows one code early (at 511, 1023, 2047).
bool lzw_decode(const unsigned char *src, size_t n, unsigned char *dst, size_t want) {
  if (n >= 2 && src[0] == 0 && (src[1] & 1)) return false;  // pre-6.0 bit order, not written since 1992
  std::vector<uint16_t> prefix(4096);
  std::vector<unsigned char> suffix(4096), first(4096);
  std::vector<uint16_t> len(4096);
  for (int i = 0; i < 256; i++) { prefix[i] = 0xffff; suffix[i] = first[i] = (unsigned char)i; len[i] = 1; }
  size_t bitpos = 0, out = 0;
  const size_t nbits = n * 8;
  int width = 9, next = 258, old = -1;
  auto get = [&](int w) -> int {
    if (bitpos + w > nbits) return 257;  // running out of data ends the strip
    int c = 0;
    for (int k = 0; k < w; k++, bitpos++) c = (c << 1) | ((src[bitpos >> 3] >> (7 - (bitpos & 7))) & 1);
    return c;
  };
  auto emit = [&](int code) -> bool {
    const int l = len[code];
    if (out + l > want) return false;
    int c = code;
    for (int k = l - 1; k >= 0; k--) { dst[out + k] = suffix[c]; c = prefix[c]; }
    out += l;
    return true;
  };
  while (out < want) {
    int code = get(width);
    if (code == 257) break;
    if (code == 256) {
      width = 9;
      next = 258;
      code = get(width);
      if (code == 257) break;
      if (code > 255) return false;
      if (!emit(code)) return false;
      old = code;
      continue;
    }
    if (old < 0) return false;  // the first code after the start must follow a clear
    int add_first;
    bool pending = false;  // the code is the entry about to be added (string(old) + its first byte)
    if (code < next) {
      if (!emit(code)) return false;
      add_first = first[code];
    } else if (code == next && next < 4096) {
      add_first = first[old];
      pending = true;
    } else {
      return false;
    }
    if (next < 4096) {
      prefix[next] = (uint16_t)old;
      suffix[next] = (unsigned char)add_first;
      first[next] = first[old];
      len[next] = (uint16_t)(len[old] + 1);
      next++;
    }
    if (pending && !emit(code)) return false;
    old = code;
    if (next + 1 > (1 << width) - 1 && width < 12) width++;
  }
  return out == want;
}

bool packbits_decode(const unsigned char *src, size_t n, unsigned char *dst, size_t want) {
  size_t i = 0, out = 0;
  while (out < want && i < n) {
    const int c = (signed char)src[i++];
    if (c >= 0) {
      const size_t k = (size_t)c + 1;
      if (i + k > n || out + k > want) return false;
      memcpy(dst + out, src + i, k);
      i += k;
      out += k;
    } else if (c != -128) {
      const size_t k = (size_t)(1 - c);
      if (i >= n || out + k > want) return false;
      memset(dst + out, src[i++], k);
      out += k;
    }
  }
  return out == want;
}

bool deflate_decode(const unsigned char *src, size_t n, unsigned char *dst, size_t want) {
  z_stream z;
  memset(&z, 0, sizeof z);
  if (inflateInit(&z) != Z_OK) return false;
  z.next_in = const_cast<unsigned char *>(src);
  z.avail_in = (uInt)n;
  z.next_out = dst;
  z.avail_out = (uInt)want;
  const int r = inflate(&z, Z_FINISH);
  const size_t got = want - z.avail_out;
  inflateEnd(&z);
  return (r == Z_STREAM_END || r == Z_OK || r == Z_BUF_ERROR) && got == want;
}

}  // namespace

extern "C" {

const char *wndgpu_io_error(void) { return g_err.c_str(); }

void wndgpu_free(void *p) { free(p); }

int wndgpu_tiff_read(const char *path, void **pixels, wndgpu_image *image) {
  if (!path || !pixels || !image) return io_fail(WNDGPU_E_ARG, "null argument");
  *pixels = nullptr;
  std::vector<unsigned char> file;
  if (!read_file(path, file)) return io_fail(WNDGPU_E_IO, "cannot read '%s': %s", path, strerror(errno));
  Tiff t{file.data(), file.size(), false};
  if (t.n < 8 || !((file[0] == 'I' && file[1] == 'I') || (file[0] == 'M' && file[1] == 'M')))
    return io_fail(WNDGPU_E_FORMAT, "'%s' is not a TIFF file", path);
  t.be = file[0] == 'M';
  if (t.u16(2) != 42) return io_fail(WNDGPU_E_FORMAT, "'%s': not a classic TIFF (BigTIFF is not read)", path);
  const uint32_t ifd = t.u32(4);
  if (!t.in(ifd, 2)) return io_fail(WNDGPU_E_FORMAT, "'%s': directory offset out of the file", path);
  const unsigned ne = t.u16(ifd);
  if (!t.in(ifd + 2, (size_t)ne * 12 + 4)) return io_fail(WNDGPU_E_FORMAT, "'%s': truncated directory", path);
  std::vector<std::pair<unsigned, Field>> tags;
  for (unsigned k = 0; k < ne; k++) {
    const size_t e = ifd + 2 + 12 * (size_t)k;
    Field fd;
    fd.type = t.u16(e + 2);
    fd.count = t.u32(e + 4);
    const size_t sz = fd.type == 3 ? 2 : fd.type == 4 ? 4 : fd.type == 1 ? 1 : 8;
    fd.off = (uint64_t)sz * fd.count <= 4 ? e + 8 : t.u32(e + 8);
    tags.emplace_back(t.u16(e), fd);
  }
  // LoadTIFF reads one image: a second directory makes it fail (cmatrix.cpp:124)
  if (t.u32(ifd + 2 + 12 * (size_t)ne) != 0)
    return io_fail(WNDGPU_E_FORMAT, "'%s' has more than one directory (image plane)", path);
  auto get = [&](unsigned tag, uint32_t dflt, bool *present = nullptr) -> uint32_t {
    for (auto &p : tags)
      if (p.first == tag) {
        std::vector<uint32_t> v;
        if (present) *present = true;
        if (field_values(t, p.second, v) && !v.empty()) return v[0];
        return dflt;
      }
    if (present) *present = false;
    return dflt;
  };
  auto get_all = [&](unsigned tag, std::vector<uint32_t> &v) -> bool {
    for (auto &p : tags)
      if (p.first == tag) return field_values(t, p.second, v);
    return false;
  };
  const uint32_t w = get(256, 0), h = get(257, 0);
  const uint32_t bps = get(258, 1), comp = get(259, 1), fill = get(266, 1), planar = get(284, 1);
  const uint32_t pred = get(317, 1);
  uint32_t spp = get(277, 1);
  if (!spp) spp = 1;  // "assume one sample per pixel if nothing is specified"
  bool tiled = false;
  get(322, 0, &tiled);
  if (!w || !h) return io_fail(WNDGPU_E_FORMAT, "'%s': missing image size", path);
  if (bps != 8 && bps != 16)
    return io_fail(WNDGPU_E_FORMAT, "'%s': %u bits per sample (only 8 and 16 are read)", path, bps);
  if (spp != 1 && spp != 3)
    return io_fail(WNDGPU_E_FORMAT, "'%s': %u samples per pixel (only 1 and 3 are read)", path, spp);
  if (tiled) return io_fail(WNDGPU_E_FORMAT, "'%s': tiled TIFF is not read", path);
  if (planar != 1 && spp > 1) return io_fail(WNDGPU_E_FORMAT, "'%s': separate colour planes are not read", path);
  if (fill != 1) return io_fail(WNDGPU_E_FORMAT, "'%s': fill order %u is not read", path, fill);
  if (pred != 1 && pred != 2) return io_fail(WNDGPU_E_FORMAT, "'%s': predictor %u is not read", path, pred);
  if (comp != 1 && comp != 5 && comp != 8 && comp != 32946 && comp != 32773)
    return io_fail(WNDGPU_E_FORMAT, "'%s': compression %u is not read", path, comp);
  if (w > 1u << 20 || h > 1u << 20) return io_fail(WNDGPU_E_FORMAT, "'%s': image too large", path);
  // RowsPerStrip absent or 0: the whole image is one strip (the TIFF default 2^32-1)
  uint32_t rps = std::min<uint32_t>(get(278, h), h);
  if (!rps) rps = h;
  std::vector<uint32_t> offs, counts;
  if (!get_all(273, offs) || offs.empty()) return io_fail(WNDGPU_E_FORMAT, "'%s': no strip offsets", path);
  const bool have_counts = get_all(279, counts);
  const size_t nstrips = ((size_t)h + rps - 1) / rps;
  if (offs.size() < nstrips) return io_fail(WNDGPU_E_FORMAT, "'%s': %zu strips for %u rows", path, offs.size(), h);
  const size_t bytes_ps = bps / 8;
  const size_t line = (size_t)w * spp * bytes_ps;
  const size_t total = line * h;
  unsigned char *buf = (unsigned char *)malloc(total ? total : 1);
  if (!buf) return io_fail(WNDGPU_E_NOMEM, "out of host memory for '%s'", path);
  for (size_t s = 0; s < nstrips; s++) {
    const size_t row0 = s * rps;
    const size_t rows = std::min<size_t>(rps, h - row0);
    const size_t want = rows * line;
    const size_t off = offs[s];
    size_t cnt = have_counts && s < counts.size() ? counts[s] : (comp == 1 ? want : t.n - std::min(off, t.n));
    if (comp == 1) cnt = std::max(cnt, want);
    if (!t.in(off, comp == 1 ? want : std::min(cnt, t.n - std::min(off, t.n)))) {
      free(buf);
      return io_fail(WNDGPU_E_FORMAT, "'%s': strip %zu lies outside the file", path, s);
    }
    cnt = std::min(cnt, t.n - off);
    unsigned char *dst = buf + row0 * line;
    bool ok;
    if (comp == 1) { memcpy(dst, t.d + off, want); ok = true; }
    else if (comp == 5) ok = lzw_decode(t.d + off, cnt, dst, want);
    else if (comp == 32773) ok = packbits_decode(t.d + off, cnt, dst, want);
    else ok = deflate_decode(t.d + off, cnt, dst, want);
    if (!ok) {
      free(buf);
      return io_fail(WNDGPU_E_FORMAT, "'%s': strip %zu does not decode to %zu bytes", path, s, want);
    }
  }
  // 16-bit samples to host order, then the horizontal predictor per row
  if (bytes_ps == 2 && t.be)
    for (size_t i = 0; i < total; i += 2) std::swap(buf[i], buf[i + 1]);
  if (pred == 2) {
    for (size_t y = 0; y < h; y++) {
      if (bytes_ps == 1) {
        unsigned char *r = buf + y * line;
        for (size_t i = spp; i < (size_t)w * spp; i++) r[i] = (unsigned char)(r[i] + r[i - spp]);
      } else {
        uint16_t *r = (uint16_t *)(buf + y * line);
        for (size_t i = spp; i < (size_t)w * spp; i++) r[i] = (uint16_t)(r[i] + r[i - spp]);
      }
    }
  }
  *pixels = buf;
  image->data = buf;
  image->height = (int)h;
  image->width = (int)w;
  image->stride = (int)w;
  image->dtype = spp == 3 ? (bps == 8 ? WNDGPU_RGB8 : WNDGPU_RGB16) : (bps == 8 ? WNDGPU_U8 : WNDGPU_U16);
  return WNDGPU_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ text formats
namespace {

// a line without its trailing '\n' / '\r' characters (chomp, TrainingSet.cpp:3042-3049);
// false at the end of the file
bool get_line(FILE *f, std::string &s) {
  s.clear();
  int c;
  bool any = false;
  while ((c = fgetc(f)) != EOF) {
    any = true;
    if (c == '\n') break;
    s.push_back((char)c);
  }
  while (!s.empty() && (s.back() == '\n' || s.back() == '\r')) s.pop_back();
  return any;
}

// "%.8g", or "%ld" for integer values (TrainingSet::SaveToFile): the
// reference tests v == (int)v, which on x86 holds for integers in the int range
void put_fit_value(FILE *f, double v) {
  if (v >= -2147483648.0 && v <= 2147483647.0 && v == (double)(int)v) fprintf(f, "%ld ", (long)v);
  else fprintf(f, "%.8g ", v);
}

}  // namespace

struct wndgpu_table {
  int version = 1, vec_type = 0, n_classes = 0;
  int n_samples = 0, n_features = 0;
  std::vector<double> values;          // [n_samples][n_features]
  std::vector<int> classes;            // [n_samples]
  std::vector<double> sample_values;   // [n_samples]
  std::vector<std::string> names, labels, paths;
};

extern "C" {

int wndgpu_sig_write(const char *path, const char *image_path, int sample_class, const double *sample_value,
                     int version, int vec_type, const double *values, const char *const *names, int n) {
  if (!path || !image_path || (!values && n > 0) || n < 0) return io_fail(WNDGPU_E_ARG, "null argument");
  FILE *f = fopen(path, "w");
  if (!f) return io_fail(WNDGPU_E_IO, "cannot write '%s': %s", path, strerror(errno));
  if (sample_value) fprintf(f, "%f\t%d.%d\n", *sample_value, version, vec_type);
  else fprintf(f, "%d\t%d.%d\n", sample_class, version, vec_type);
  fprintf(f, "%s\n", image_path);
  for (int i = 0; i < n; i++) {
    if (names) fprintf(f, "%.8g\t%s\n", values[i], names[i]);
    else fprintf(f, "%.8g\n", values[i]);
  }
  const bool ok = !ferror(f);
  if (fclose(f) != 0 || !ok) return io_fail(WNDGPU_E_IO, "write error on '%s'", path);
  return WNDGPU_OK;
}

int wndgpu_sig_read(const char *path, int continuous, wndgpu_table **out) {
  if (!path || !out) return io_fail(WNDGPU_E_ARG, "null argument");
  *out = nullptr;
  FILE *f = fopen(path, "r");
  if (!f) return io_fail(WNDGPU_E_IO, "cannot read '%s': %s", path, strerror(errno));
  wndgpu_table *t = new wndgpu_table;
  std::string s;
  get_line(f, s);
  // class (or value) and version "maj.min"; no version means 1.0
  int vmaj = 0, vmin = 0;
  double sv = 0;
  unsigned short cls = 0;
  if (continuous) {
    sscanf(s.c_str(), "%lf%*[\t ]%d.%d", &sv, &vmaj, &vmin);
    cls = 1;
  } else {
    sscanf(s.c_str(), "%hu%*[\t ]%d.%d", &cls, &vmaj, &vmin);
    sv = cls;
  }
  t->version = vmaj ? vmaj : 1;
  t->vec_type = vmaj ? vmin : 0;
  get_line(f, s);
  t->paths.push_back(s);
  // "value<tab>name" per line; a line without a number repeats the previous
  // value (the reference's sscanf leaves it in place)
  double val = 0;
  while (get_line(f, s)) {
    const char *p = s.c_str();
    char *end = nullptr;
    const double v = strtod(p, &end);
    std::string name;
    if (end != p) {
      val = v;
      if (*end == '\t' || *end == ' ') {
        while (*end == '\t' || *end == ' ') end++;
        const char *q = end;
        while (*q && *q != '\t' && *q != '\r' && *q != '\n') q++;
        name.assign((const char *)end, q);
      }
    }
    t->values.push_back(val);
    t->names.push_back(name);
  }
  fclose(f);
  t->n_samples = 1;
  t->n_features = (int)t->values.size();
  t->classes.push_back(cls);
  t->sample_values.push_back(sv);
  *out = t;
  return WNDGPU_OK;
}

int wndgpu_fit_write(const char *path, int n_classes, int version, int vec_type, int n_samples, int n_features,
                     const double *values, const int *classes, const double *sample_values,
                     const char *const *feature_names, const char *const *class_labels,
                     const char *const *sample_paths) {
  if (!path || !feature_names || !class_labels || n_samples < 0 || n_features < 0 || n_classes < 0 ||
      (n_samples && (!values || (!classes && !sample_values))))
    return io_fail(WNDGPU_E_ARG, "null argument");
  FILE *f = fopen(path, "w");
  if (!f) return io_fail(WNDGPU_E_IO, "cannot write '%s': %s", path, strerror(errno));
  fprintf(f, "%ld\t%d.%d\n", (long)n_classes, version, vec_type);
  fprintf(f, "%ld\n", (long)n_features);
  fprintf(f, "%ld\n", (long)n_samples);
  for (int i = 0; i < n_features; i++) fprintf(f, "%s\n", feature_names[i]);
  for (int c = 0; c <= n_classes; c++) fprintf(f, "%s\n", class_labels[c]);
  for (int s = 0; s < n_samples; s++) {
    for (int i = 0; i < n_features; i++) put_fit_value(f, values[(size_t)s * n_features + i]);
    if (sample_values) fprintf(f, "%f\n", sample_values[s]);
    else fprintf(f, "%d\n", classes[s]);
    fprintf(f, "%s\n", sample_paths ? sample_paths[s] : "");
  }
  const bool ok = !ferror(f);
  if (fclose(f) != 0 || !ok) return io_fail(WNDGPU_E_IO, "write error on '%s'", path);
  return WNDGPU_OK;
}

int wndgpu_fit_read(const char *path, wndgpu_table **out) {
  if (!path || !out) return io_fail(WNDGPU_E_ARG, "null argument");
  *out = nullptr;
  FILE *f = fopen(path, "r");
  if (!f) return io_fail(WNDGPU_E_IO, "cannot read '%s': %s", path, strerror(errno));
  wndgpu_table *t = new wndgpu_table;
  auto bad = [&](const char *what) {
    fclose(f);
    delete t;
    return io_fail(WNDGPU_E_FORMAT, "'%s': %s", path, what);
  };
  std::string s;
  if (!get_line(f, s)) return bad("empty file");
  int ncls = 0, vmaj = 0, vmin = 0;
  sscanf(s.c_str(), "%d%*[\t ]%d.%d", &ncls, &vmaj, &vmin);
  t->version = vmaj ? vmaj : 1;
  t->vec_type = vmaj ? vmin : 0;
  t->n_classes = ncls;
  get_line(f, s);
  t->n_features = atoi(s.c_str());
  get_line(f, s);
  t->n_samples = atoi(s.c_str());
  if (ncls < 0 || t->n_features < 0 || t->n_samples < 0) return bad("negative count in the header");
  for (int i = 0; i < t->n_features; i++) {
    if (!get_line(f, s)) return bad("truncated feature names");
    t->names.push_back(s);
  }
  for (int c = 0; c <= ncls; c++) {
    if (!get_line(f, s)) return bad("truncated class labels");
    t->labels.push_back(s);
  }
  t->values.resize((size_t)t->n_samples * t->n_features);
  for (int k = 0; k < t->n_samples; k++) {
    if (!get_line(f, s)) return bad("truncated samples");
    // space-separated values, then the class (strtok(" \n") + atof / atoi)
    std::vector<char> b(s.begin(), s.end());
    b.push_back('\0');
    char *save = nullptr;
    char *tok = strtok_r(b.data(), " \n", &save);
    for (int i = 0; i < t->n_features; i++) {
      if (!tok) return bad("a sample line has fewer values than features");
      t->values[(size_t)k * t->n_features + i] = atof(tok);
      tok = strtok_r(nullptr, " \n", &save);
    }
    if (!tok) return bad("a sample line has no class");
    const int cls = atoi(tok);
    t->classes.push_back(cls);
    t->sample_values.push_back(atof(tok));  // the value of a continuous set's sample
    get_line(f, s);
    t->paths.push_back(s);
  }
  fclose(f);
  *out = t;
  return WNDGPU_OK;
}

int wndgpu_table_shape(const wndgpu_table *t, int *n_samples, int *n_features, int *n_classes, int *version,
                       int *vec_type) {
  if (!t) return io_fail(WNDGPU_E_ARG, "null table");
  if (n_samples) *n_samples = t->n_samples;
  if (n_features) *n_features = t->n_features;
  if (n_classes) *n_classes = t->n_classes;
  if (version) *version = t->version;
  if (vec_type) *vec_type = t->vec_type;
  return WNDGPU_OK;
}
const double *wndgpu_table_values(const wndgpu_table *t) { return t ? t->values.data() : nullptr; }
const int *wndgpu_table_classes(const wndgpu_table *t) { return t ? t->classes.data() : nullptr; }
const double *wndgpu_table_sample_values(const wndgpu_table *t) { return t ? t->sample_values.data() : nullptr; }
const char *wndgpu_table_feature_name(const wndgpu_table *t, int i) {
  return t && i >= 0 && i < (int)t->names.size() ? t->names[i].c_str() : nullptr;
}
const char *wndgpu_table_class_label(const wndgpu_table *t, int i) {
  return t && i >= 0 && i < (int)t->labels.size() ? t->labels[i].c_str() : nullptr;
}
const char *wndgpu_table_sample_path(const wndgpu_table *t, int i) {
  return t && i >= 0 && i < (int)t->paths.size() ? t->paths[i].c_str() : nullptr;
}
void wndgpu_table_destroy(wndgpu_table *t) { delete t; }

int wndgpu_weights_write(const char *path, const double *weights, const char *const *names, int n) {
  if (!path || (!weights && n > 0) || (!names && n > 0) || n < 0) return io_fail(WNDGPU_E_ARG, "null argument");
  FILE *f = fopen(path, "w");
  if (!f) return io_fail(WNDGPU_E_IO, "cannot write '%s': %s", path, strerror(errno));
  for (int i = 0; i < n; i++) fprintf(f, "%f %s\n", weights[i], names[i]);
  const bool ok = !ferror(f);
  if (fclose(f) != 0 || !ok) return io_fail(WNDGPU_E_IO, "write error on '%s'", path);
  return WNDGPU_OK;
}

int wndgpu_weights_read(const char *path, double *weights, int cap, int *n_read) {
  if (!path || !n_read || (cap > 0 && !weights)) return io_fail(WNDGPU_E_ARG, "null argument");
  FILE *f = fopen(path, "r");
  if (!f) return io_fail(WNDGPU_E_IO, "cannot read '%s': %s", path, strerror(errno));
  // 128-byte reads as the reference's fgets: every read is one weight, the
  // number before the first space; negative weights become 0
  char line[128];
  int n = 0;
  while (fgets(line, sizeof line, f)) {
    if (!line[0]) continue;
    if (char *sp = strchr(line, ' ')) *sp = '\0';
    double v = atof(line);
    if (v < 0) v = 0;
    if (n < cap) weights[n] = v;
    n++;
  }
  fclose(f);
  *n_read = n;
  if (n > cap) return io_fail(WNDGPU_E_ARG, "'%s' holds %d weights, more than %d", path, n, cap);
  return WNDGPU_OK;
}

}  // extern "C"
