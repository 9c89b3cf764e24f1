// Mock MEX runtime for testing the engine's MEX gateway (mrilab_amd/mex/)
// without MATLAB. TEST INFRASTRUCTURE ONLY.
//
// It keeps a "global workspace" of one-element structs whose fields are
// typed, column-major arrays. A test fills it through mock_set (a fresh
// buffer every time, so replacing a field moves its data as MATLAB's
// reallocation does), runs mexFunction through mock_run, and reads fields
// back through mock_get. mexEvalString hands the command ("DoExtPlugin",
// "DoGPUFetch") to a handler the test registers; mexErrMsgIdAndTxt ends the
// MEX call with the message, as MATLAB's longjmp does.
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "mex.h"

struct mxArray_tag {
    mxClassID cls = mxUNKNOWN_CLASS;
    std::vector<mwSize> dims;
    std::vector<unsigned char> bytes;
    std::map<std::string, std::unique_ptr<mxArray_tag>> fields;  // struct arrays
};

namespace {

struct MexError {
    std::string id, msg;
};

std::map<std::string, std::unique_ptr<mxArray_tag>> g_globals;
std::vector<std::unique_ptr<mxArray_tag>> g_retired;  // replaced arrays stay valid until mock_reset
int (*g_eval)(const char *) = nullptr;
std::string g_err, g_err_id;
int g_unshare = 0;

size_t class_size(mxClassID c) {
    switch (c) {
        case mxDOUBLE_CLASS: return 8;
        case mxSINGLE_CLASS: return 4;
        case mxINT32_CLASS: return 4;
        case mxCHAR_CLASS: return 2;  // MATLAB chars are 16-bit
        default: return 1;
    }
}

size_t numel(const mxArray_tag *a) {
    size_t n = 1;
    for (mwSize d : a->dims) n *= d;
    return a->dims.empty() ? 0 : n;
}

}  // namespace

extern "C" {

const mxArray *mexGetVariablePtr(const char *workspace, const char *name) {
    if (std::strcmp(workspace, "global") != 0) return nullptr;
    auto it = g_globals.find(name);
    return it == g_globals.end() ? nullptr : it->second.get();
}

mxArray *mxGetField(const mxArray *pm, mwIndex index, const char *fieldname) {
    if (!pm || index != 0) return nullptr;
    auto it = pm->fields.find(fieldname);
    return it == pm->fields.end() ? nullptr : it->second.get();
}

void *mxGetData(const mxArray *pm) { return pm && !pm->bytes.empty() ? (void *)pm->bytes.data() : nullptr; }
size_t mxGetNumberOfElements(const mxArray *pm) { return pm ? numel(pm) : 0; }
mwSize mxGetNumberOfDimensions(const mxArray *pm) { return pm ? pm->dims.size() : 0; }
const mwSize *mxGetDimensions(const mxArray *pm) { return pm ? pm->dims.data() : nullptr; }
mxClassID mxGetClassID(const mxArray *pm) { return pm ? pm->cls : mxUNKNOWN_CLASS; }

int mxGetString(const mxArray *pm, char *str, mwSize len) {
    if (!pm || pm->cls != mxCHAR_CLASS || len == 0) return 1;
    const size_t n = numel(pm);
    const uint16_t *c = (const uint16_t *)pm->bytes.data();
    size_t k = 0;
    for (; k < n && k + 1 < len; k++) str[k] = (char)c[k];
    str[k] = 0;
    return k < n ? 1 : 0;
}

int mexEvalString(const char *command) { return g_eval ? g_eval(command) : 1; }

void mexErrMsgIdAndTxt(const char *errorid, const char *errormsg, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, errormsg);
    std::vsnprintf(buf, sizeof buf, errormsg, ap);
    va_end(ap);
    throw MexError{errorid ? errorid : "", buf};
}

int mxUnshareArray(mxArray *pa, int noDeepCopy) {
    (void)pa;
    (void)noDeepCopy;
    g_unshare++;
    return 0;
}

// ---- test control ------------------------------------------------------

void mock_reset(void) {
    g_globals.clear();
    g_retired.clear();
    g_eval = nullptr;
    g_err.clear();
    g_err_id.clear();
    g_unshare = 0;
}

// Sets global var.field to a fresh copy of ndims/dims/data of class cls
// (creating the struct); the array it replaces stays readable until reset.
int mock_set(const char *var, const char *fieldname, int cls, int ndims, const int64_t *dims, const void *data) {
    auto &g = g_globals[var];
    if (!g) {
        g.reset(new mxArray_tag());
        g->cls = mxSTRUCT_CLASS;
        g->dims = {1, 1};
    }
    std::unique_ptr<mxArray_tag> a(new mxArray_tag());
    a->cls = (mxClassID)cls;
    for (int i = 0; i < ndims; i++) a->dims.push_back((mwSize)dims[i]);
    const size_t n = numel(a.get()) * class_size(a->cls);
    a->bytes.resize(n);
    if (n && data) std::memcpy(a->bytes.data(), data, n);
    auto it = g->fields.find(fieldname);
    if (it != g->fields.end()) g_retired.push_back(std::move(it->second));
    g->fields[fieldname] = std::move(a);
    return 0;
}

// Removes global var.field (returns 1 if there was none).
int mock_remove(const char *var, const char *fieldname) {
    auto g = g_globals.find(var);
    if (g == g_globals.end()) return 1;
    auto it = g->second->fields.find(fieldname);
    if (it == g->second->fields.end()) return 1;
    g_retired.push_back(std::move(it->second));
    g->second->fields.erase(it);
    return 0;
}

// Data of global var.field (NULL if absent), its class and element count.
void *mock_get(const char *var, const char *fieldname, int *cls, int64_t *n) {
    auto g = g_globals.find(var);
    if (g == g_globals.end()) return nullptr;
    auto it = g->second->fields.find(fieldname);
    if (it == g->second->fields.end()) return nullptr;
    if (cls) *cls = (int)it->second->cls;
    if (n) *n = (int64_t)numel(it->second.get());
    return it->second->bytes.empty() ? nullptr : it->second->bytes.data();
}

void mock_set_eval(int (*fn)(const char *)) { g_eval = fn; }

// Runs mexFunction as MATLAB does (no arguments); 0 on success, 1 when it
// raised an error (mock_error / mock_error_id).
int mock_run(void) {
    g_err.clear();
    g_err_id.clear();
    try {
        mexFunction(0, nullptr, 0, nullptr);
    } catch (const MexError &e) {
        g_err = e.msg;
        g_err_id = e.id;
        return 1;
    }
    return 0;
}

const char *mock_error(void) { return g_err.c_str(); }
const char *mock_error_id(void) { return g_err_id.c_str(); }
int mock_unshare_count(void) { return g_unshare; }

}  // extern "C"
