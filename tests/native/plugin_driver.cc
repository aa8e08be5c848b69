// C entry points over the "frsz2" LibPressio plugin (frsz_amd/plugin/frsz2_plugin.cc)
// built against the LibPressio test double (tests/pressio_double), for the
// ctypes tests tests/test_plugin.py and tests/test_gpu_plugin.py.  The plugin
// is looked up in the compressor registry by name, as a LibPressio client
// does.  Test infrastructure.
#include "frsz.h"
#include "libpressio_ext/cpp/pressio.h"

#include <cstdint>
#include <cstring>
#include <memory>
#include <string>

namespace {

struct handle
{
  std::shared_ptr<libpressio_compressor_plugin> p;
  pressio_data compressed;
};

pressio_dtype
dtype_of(int d)
{
  return d == 0 ? pressio_float_dtype : d == 1 ? pressio_double_dtype : pressio_int32_dtype;
}

int
copy_text(std::string const& s, char* buf, int len)
{
  if (len <= 0) {
    return -1;
  }
  std::strncpy(buf, s.c_str(), static_cast<std::size_t>(len - 1));
  buf[len - 1] = '\0';
  return 0;
}

} // namespace

extern "C" {

void*
pd_create(const char* name)
{
  register_frsz();
  auto p = compressor_plugins().build(name);
  if (!p) {
    return nullptr;
  }
  return new handle{ std::shared_ptr<libpressio_compressor_plugin>(std::move(p)), {} };
}

void*
pd_clone(void* h)
{
  return new handle{ static_cast<handle*>(h)->p->clone(), {} };
}

void
pd_free(void* h)
{
  delete static_cast<handle*>(h);
}

int
pd_set_u64(void* h, const char* key, std::uint64_t v)
{
  pressio_options o;
  pressio_options_set(o, key, v);
  return static_cast<handle*>(h)->p->set_options(o);
}

int
pd_set_i32(void* h, const char* key, std::int32_t v)
{
  pressio_options o;
  pressio_options_set(o, key, v);
  return static_cast<handle*>(h)->p->set_options(o);
}

// 0 found, 1 absent, 2 not a number
int
pd_get_u64(void* h, const char* key, std::uint64_t* v)
{
  return pressio_options_get(static_cast<handle*>(h)->p->get_options(), key, v);
}

int
pd_get_config_i64(void* h, const char* key, std::int64_t* v)
{
  return pressio_options_get(static_cast<handle*>(h)->p->get_configuration(), key, v);
}

int
pd_get_config_str(void* h, const char* key, char* buf, int len)
{
  std::string s;
  const int rc = pressio_options_get(static_cast<handle*>(h)->p->get_configuration(), key, &s);
  return rc ? rc : copy_text(s, buf, len);
}

int
pd_get_doc_str(void* h, const char* key, char* buf, int len)
{
  std::string s;
  const int rc = pressio_options_get(static_cast<handle*>(h)->p->get_documentation(), key, &s);
  return rc ? rc : copy_text(s, buf, len);
}

int
pd_version(void* h, int* major, int* minor, int* patch, char* buf, int len)
{
  auto& p = *static_cast<handle*>(h)->p;
  *major = p.major_version();
  *minor = p.minor_version();
  *patch = p.patch_version();
  return copy_text(std::string(p.version()) + " " + p.prefix(), buf, len);
}

// compress n host values of dtype (0 float, 1 double, 2 int32) held at `in`;
// the stream stays in the handle (pd_stream_bytes / pd_copy_stream)
int
pd_compress(void* h, int dtype, void* in, std::uint64_t n)
{
  auto* H = static_cast<handle*>(h);
  const pressio_data input = pressio_data::nonowning(dtype_of(dtype), in, { static_cast<std::size_t>(n) });
  return H->p->compress(&input, &H->compressed);
}

std::uint64_t
pd_stream_bytes(void* h)
{
  return static_cast<handle*>(h)->compressed.size_in_bytes();
}

int
pd_copy_stream(void* h, void* dst, std::uint64_t cap)
{
  const pressio_data& c = static_cast<handle*>(h)->compressed;
  if (cap < c.size_in_bytes() || !c.data()) {
    return -1;
  }
  std::memcpy(dst, c.data(), c.size_in_bytes());
  return 0;
}

// decompress in_bytes of stream at `in` into n values of dtype at `out`
// (the output's dtype and element count drive the decode)
int
pd_decompress(void* h, void* in, std::uint64_t in_bytes, int dtype, void* out, std::uint64_t n)
{
  const pressio_data input = pressio_data::nonowning(pressio_byte_dtype, in, { static_cast<std::size_t>(in_bytes) });
  pressio_data output = pressio_data::nonowning(dtype_of(dtype), out, { static_cast<std::size_t>(n) });
  return static_cast<handle*>(h)->p->decompress(&input, &output);
}

int
pd_error_code(void* h)
{
  return static_cast<handle*>(h)->p->error_code();
}

const char*
pd_error_msg(void* h)
{
  return static_cast<handle*>(h)->p->error_msg();
}

} // extern "C"
