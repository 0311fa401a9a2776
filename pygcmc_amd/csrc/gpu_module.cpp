// gpu_module.cpp -- CPython module pygcmc_amd.gpu, the drop-in for pyGCMC's
// gcmc.gpu extension (gcmc/cpp/gcmc.cpp:162-225).
//
// runGCMC(info, fragmentInfo, residueInfo, atomInfo, grid, ff, moveArray)
// takes the same seven numpy structured arrays (dtypes of
// gcmc/python/values.py:18-63), runs info[0].mcsteps moves on the GPU through
// the C ABI (runGCMC_hip) and updates fragmentInfo[].totalNum, residueInfo and
// atomInfo in place, returning None.  Unlike the reference it checks item
// sizes, contiguity and lengths (TypeError / ValueError), releases the GIL
// while the GPU runs, and raises RuntimeError on a device failure.
#define PY_SSIZE_T_CLEAN
#include <Python.h>

#include <cstdio>

#include "pygcmc_hip.h"

namespace {

struct View {
    Py_buffer b{};
    bool ok = false;
    ~View()
    {
        if (ok) PyBuffer_Release(&b);
    }
};

bool get_view(PyObject *o, View &v, const char *name, Py_ssize_t itemsize, bool writable)
{
    int flags = PyBUF_C_CONTIGUOUS | PyBUF_FORMAT | (writable ? PyBUF_WRITABLE : 0);
    if (PyObject_GetBuffer(o, &v.b, flags) != 0) {
        PyErr_Clear();
        PyErr_Format(PyExc_TypeError, "%s must be a C-contiguous%s array", name,
                     writable ? ", writable" : "");
        return false;
    }
    v.ok = true;
    if (v.b.itemsize != itemsize) {
        PyErr_Format(PyExc_TypeError, "%s has item size %zd, expected %zd", name, v.b.itemsize, itemsize);
        return false;
    }
    return true;
}

Py_ssize_t count(const View &v) { return v.b.len / v.b.itemsize; }

bool need(const View &v, const char *name, Py_ssize_t n)
{
    if (count(v) < n) {
        PyErr_Format(PyExc_ValueError, "%s has %zd items, the InfoStruct needs %zd", name, count(v), n);
        return false;
    }
    return true;
}

PyObject *runGCMC(PyObject *, PyObject *args)
{
    PyObject *o[7];
    if (!PyArg_ParseTuple(args, "OOOOOOO", &o[0], &o[1], &o[2], &o[3], &o[4], &o[5], &o[6])) return nullptr;
    View vi, vf, vr, va, vg, vff, vm;
    if (!get_view(o[0], vi, "info", sizeof(InfoStruct), false)) return nullptr;
    if (!get_view(o[1], vf, "fragmentInfo", sizeof(AtomArray), true)) return nullptr;
    if (!get_view(o[2], vr, "residueInfo", sizeof(residue), true)) return nullptr;
    if (!get_view(o[3], va, "atomInfo", sizeof(Atom), true)) return nullptr;
    if (!get_view(o[4], vg, "grid", sizeof(float), false)) return nullptr;
    if (!get_view(o[5], vff, "ff", sizeof(float), false)) return nullptr;
    if (!get_view(o[6], vm, "moveArray", sizeof(int), false)) return nullptr;
    if (count(vi) < 1) {
        PyErr_SetString(PyExc_ValueError, "info is empty");
        return nullptr;
    }
    const InfoStruct *info = (const InfoStruct *)vi.b.buf;
    if (info->fragTypeNum < 1 || info->totalResNum < 0 || info->totalAtomNum < 0 || info->totalGridNum < 0 ||
        info->ffXNum < 0 || info->ffYNum < 0 || info->mcsteps < 0) {
        PyErr_SetString(PyExc_ValueError, "info holds negative counts");
        return nullptr;
    }
    if (!need(vf, "fragmentInfo", info->fragTypeNum) || !need(vr, "residueInfo", info->totalResNum) ||
        !need(va, "atomInfo", info->totalAtomNum) || !need(vg, "grid", 3 * (Py_ssize_t)info->totalGridNum) ||
        !need(vff, "ff", 2 * (Py_ssize_t)info->ffXNum * info->ffYNum) || !need(vm, "moveArray", info->mcsteps))
        return nullptr;
    if (info->showInfo > 0)
        std::printf("runGCMC: %d steps, %d species, %d residues, %d atoms, %d grid points\n", info->mcsteps,
                    info->fragTypeNum, info->totalResNum, info->totalAtomNum, info->totalGridNum);
    int rc;
    Py_BEGIN_ALLOW_THREADS
    rc = runGCMC_hip(info, (AtomArray *)vf.b.buf, (residue *)vr.b.buf, (Atom *)va.b.buf, (const float *)vg.b.buf,
                     (const float *)vff.b.buf, (const int *)vm.b.buf);
    Py_END_ALLOW_THREADS
    if (rc != 0) {
        PyErr_Format(PyExc_RuntimeError, "runGCMC failed (%d): %s", rc, gcmc_last_error());
        return nullptr;
    }
    Py_RETURN_NONE;
}

PyMethodDef methods[] = {
    {"runGCMC", runGCMC, METH_VARARGS, "Run GCMC on the GPU (MI355X HIP engine)."},
    {nullptr, nullptr, 0, nullptr},
};

PyModuleDef moduledef = {PyModuleDef_HEAD_INIT, "gpu", "GCMC hot path on AMD GPUs", -1, methods,
                         nullptr, nullptr, nullptr, nullptr};

}  // namespace

PyMODINIT_FUNC PyInit_gpu(void) { return PyModule_Create(&moduledef); }
