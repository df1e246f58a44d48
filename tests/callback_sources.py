"""Callback sources of the callback tests (reference: test/callback.cpp:41-63,139-158).

load: reads the N_ext/2+1-bin layout of a c2r plan from a buffer holding only N/2+1 bins per
column and returns zero for the missing bins (zero padding).
store: keeps the first N/4 bins of an r2c output, scaled by 1/N, in a truncated layout.
"""


def load_hip(real_t, M, n_ref, n_have):
    t2 = real_t + "2"
    return f"""
__device__ {t2} load({t2}* in, size_t offset) {{
    unsigned n = offset / {M}ul % {n_ref}ul;
    if (n < {n_have}u) {{
        unsigned long k = offset / {M * n_ref}ul;
        return in[offset - k * {M * (n_ref - n_have)}ul];
    }}
    {t2} z; z.x = 0; z.y = 0;
    return z;
}}
"""


def store_hip(real_t, M, n_ref, n_cut, scale_hex):
    t2 = real_t + "2"
    return f"""
__device__ void store({t2}* out, size_t offset, {t2} value) {{
    unsigned n = offset / {M}ul % {n_ref}ul;
    if (n < {n_cut}u) {{
        unsigned long k = offset / {M * n_ref}ul;
        {t2} v; v.x = value.x * ({real_t}){scale_hex}; v.y = value.y * ({real_t}){scale_hex};
        out[offset - k * {M * (n_ref - n_cut)}ul] = v;
    }}
}}
"""


def store_opencl(real_t, M, n_ref, n_cut, scale_hex):
    """The same store callback in the reference's OpenCL-C dialect (global pointers, uint)."""
    t2 = real_t + "2"
    return f"""
void store(global {t2}* out, size_t offset, {t2} value) {{
    uint n = offset / {M}ul % {n_ref}ul;
    if (n < {n_cut}) {{
        ulong k = offset / {M * n_ref}ul;
        {t2} v; v.x = value.x * ({real_t}){scale_hex}; v.y = value.y * ({real_t}){scale_hex};
        out[offset - k * {M * (n_ref - n_cut)}ul] = v;
    }}
}}
"""
