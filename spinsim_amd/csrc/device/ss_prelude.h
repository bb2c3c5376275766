// spinsim_amd device prelude: everything a field callback may use.
//
// This text is compiled by hipRTC in front of the user's get_field source.
// A callback has one of two forms (the reference's plugin is
// get_field(time_sample, sweep_parameters, field_sample), writing into
// field_sample in place; spinsim/__init__.py:620 and :459-469):
//
//   __device__ void get_field(double t, const double* sweep, double* field);
//   __device__ void get_field(double t, const double* sweep, double* field,
//                             SS_CONTEXT);   // when it reads samplers
//
// field[0..2] are the x, y, z angular frequencies (rad/s) and field[3] the
// spin-one quadratic shift.  Entries the callback leaves untouched are zero.
// Interpolation samplers are read with SS_SAMPLE(k, t) or by the name they
// were registered under (a macro defined after this prelude); a multi-axis
// sampler with SS_SAMPLE_AXES(k, t, out, n) (all n components from one index
// search) or SS_SAMPLE_AXIS(k, a, t) (component a).
#pragma once

#define SS_PI 3.14159265358979323846
#define SS_TAU 6.28318530717958647692
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

// One interpolation table (generate_interpolation_sampler's frozen
// (time, amplitude) pair, spinsim/__init__.py:1941), held in device memory so
// new samples need no recompilation.  amplitude_stride > 0 gives one row per
// simulation of a batch; `axes` components share the time grid, component a
// at amplitude + a * axis_stride.  Layout shared with include/spinsim_hip.h.
struct ss_table {
    const double* time;
    const double* amplitude;
    long long size;
    long long amplitude_stride;
    int order;
    int zero_boundary;
    int axes;
    int reserved;
    long long axis_stride;
};

// The interval a table was last sampled in, as the cubic
// c0 + c1 p + c2 p^2 + c3 p^3 of p = (t - t0) inv_h; inv_h = NaN when empty;
// idx is the interval's index (knots idx, idx + 1).
struct ss_interp_line {
    double t0, inv_h, c0, c1, c2, c3;
    long long idx;
};

#ifndef SS_INTERP_LINES
#define SS_INTERP_LINES 4  // tables 0..3 keep their last interval per lane
#endif

struct ss_ctx {
    const ss_table* tables;
    int sim;
    // Per-lane cache of sampled intervals: consecutive samples of one lane
    // (a coarse step, a few knots apart at most) mostly fall in the same
    // interval, which then costs no table reads and no divisions.
    mutable ss_interp_line line[SS_INTERP_LINES];
    __device__ ss_ctx() : tables(nullptr), sim(0)
    {
#pragma unroll
        for (int k = 0; k < SS_INTERP_LINES; ++k) {
            line[k].inv_h = __builtin_nan("");
            line[k].idx = -1;
        }
    }
};

// ---- float32 transcendentals for field callbacks (precision "fp32") --------
// In that mode the callback's sin(x), cos(x) and exp(x) resolve to these.
// Arguments are reduced in float64 by Cody-Waite steps (the constant split in
// two, fused multiply-adds), so no phase is rounded to float32; float32 then
// carries only the higher-order terms:
//   sin, cos: x = k h + b, h = 2 pi / 256, |b| <= h / 2, a = k h from a table
//     of float64 cos a, sin a (cos and sin of j 2 pi / 256, correctly rounded):
//       cos x = cos a (1 + cm) - sin a b (1 + sm),
//       sin x = sin a (1 + cm) + cos a b (1 + sm),
//     cm = cos b - 1 = -b^2/2 + b^4/24, sm = sin b / b - 1 = -b^2/6 + b^4/120
//     in float32 (|cm| < 7.6e-5, |sm| < 2.6e-5: their rounding is about 1e-11
//     absolute in the result; truncation below 1e-17).
//   exp: x = (64 n + j) ln2 / 64 + b, |b| <= ln2 / 128, 2^(j/64) from a table:
//     exp x = 2^n 2^(j/64) (1 + b + b^2 q), q in float32 (about 1e-12 relative).
// Arguments outside the reductions' range (and NaN) take the float64 functions.
#define SS_TRIG_H_HI 0.02454369260617026
#define SS_TRIG_H_LO 9.567553118338697e-19
#define SS_TRIG_INV_H 40.74366543152521
#define SS_EXP_H_HI 0.010830424696249145
#define SS_EXP_H_LO 3.623510646634843e-19
#define SS_EXP_INV_H 92.33248261689366

// cos(2 pi j / 256), sin(2 pi j / 256), j = 0 .. 255
__constant__ static const double ss_trig_table[512] = {
    1.0, 0.0, 0.9996988186962042, 0.024541228522912288,
    0.9987954562051724, 0.049067674327418015, 0.9972904566786902, 0.07356456359966743,
    0.9951847266721969, 0.0980171403295606, 0.99247953459871, 0.1224106751992162,
    0.989176509964781, 0.14673047445536175, 0.9852776423889412, 0.17096188876030122,
    0.9807852804032304, 0.19509032201612828, 0.9757021300385286, 0.2191012401568698,
    0.970031253194544, 0.2429801799032639, 0.9637760657954398, 0.26671275747489837,
    0.9569403357322088, 0.2902846772544624, 0.9495281805930367, 0.31368174039889146,
    0.9415440651830208, 0.33688985339222005, 0.9329927988347388, 0.35989503653498817,
    0.9238795325112867, 0.3826834323650898, 0.9142097557035307, 0.40524131400498986,
    0.9039892931234433, 0.4275550934302821, 0.8932243011955153, 0.4496113296546066,
    0.881921264348355, 0.47139673682599764, 0.8700869911087115, 0.49289819222978404,
    0.8577286100002721, 0.5141027441932218, 0.8448535652497071, 0.5349976198870973,
    0.8314696123025452, 0.5555702330196022, 0.8175848131515837, 0.5758081914178453,
    0.8032075314806449, 0.5956993044924334, 0.7883464276266062, 0.6152315905806268,
    0.773010453362737, 0.6343932841636455, 0.7572088465064846, 0.6531728429537768,
    0.7409511253549591, 0.6715589548470184, 0.7242470829514669, 0.6895405447370669,
    0.7071067811865476, 0.7071067811865476, 0.6895405447370669, 0.7242470829514669,
    0.6715589548470184, 0.7409511253549591, 0.6531728429537768, 0.7572088465064846,
    0.6343932841636455, 0.773010453362737, 0.6152315905806268, 0.7883464276266062,
    0.5956993044924334, 0.8032075314806449, 0.5758081914178453, 0.8175848131515837,
    0.5555702330196022, 0.8314696123025452, 0.5349976198870973, 0.8448535652497071,
    0.5141027441932218, 0.8577286100002721, 0.49289819222978404, 0.8700869911087115,
    0.47139673682599764, 0.881921264348355, 0.4496113296546066, 0.8932243011955153,
    0.4275550934302821, 0.9039892931234433, 0.40524131400498986, 0.9142097557035307,
    0.3826834323650898, 0.9238795325112867, 0.35989503653498817, 0.9329927988347388,
    0.33688985339222005, 0.9415440651830208, 0.31368174039889146, 0.9495281805930367,
    0.2902846772544624, 0.9569403357322088, 0.26671275747489837, 0.9637760657954398,
    0.2429801799032639, 0.970031253194544, 0.2191012401568698, 0.9757021300385286,
    0.19509032201612828, 0.9807852804032304, 0.17096188876030122, 0.9852776423889412,
    0.14673047445536175, 0.989176509964781, 0.1224106751992162, 0.99247953459871,
    0.0980171403295606, 0.9951847266721969, 0.07356456359966743, 0.9972904566786902,
    0.049067674327418015, 0.9987954562051724, 0.024541228522912288, 0.9996988186962042,
    2.0670321098263988e-43, 1.0, -0.024541228522912288, 0.9996988186962042,
    -0.049067674327418015, 0.9987954562051724, -0.07356456359966743, 0.9972904566786902,
    -0.0980171403295606, 0.9951847266721969, -0.1224106751992162, 0.99247953459871,
    -0.14673047445536175, 0.989176509964781, -0.17096188876030122, 0.9852776423889412,
    -0.19509032201612828, 0.9807852804032304, -0.2191012401568698, 0.9757021300385286,
    -0.2429801799032639, 0.970031253194544, -0.26671275747489837, 0.9637760657954398,
    -0.2902846772544624, 0.9569403357322088, -0.31368174039889146, 0.9495281805930367,
    -0.33688985339222005, 0.9415440651830208, -0.35989503653498817, 0.9329927988347388,
    -0.3826834323650898, 0.9238795325112867, -0.40524131400498986, 0.9142097557035307,
    -0.4275550934302821, 0.9039892931234433, -0.4496113296546066, 0.8932243011955153,
    -0.47139673682599764, 0.881921264348355, -0.49289819222978404, 0.8700869911087115,
    -0.5141027441932218, 0.8577286100002721, -0.5349976198870973, 0.8448535652497071,
    -0.5555702330196022, 0.8314696123025452, -0.5758081914178453, 0.8175848131515837,
    -0.5956993044924334, 0.8032075314806449, -0.6152315905806268, 0.7883464276266062,
    -0.6343932841636455, 0.773010453362737, -0.6531728429537768, 0.7572088465064846,
    -0.6715589548470184, 0.7409511253549591, -0.6895405447370669, 0.7242470829514669,
    -0.7071067811865476, 0.7071067811865476, -0.7242470829514669, 0.6895405447370669,
    -0.7409511253549591, 0.6715589548470184, -0.7572088465064846, 0.6531728429537768,
    -0.773010453362737, 0.6343932841636455, -0.7883464276266062, 0.6152315905806268,
    -0.8032075314806449, 0.5956993044924334, -0.8175848131515837, 0.5758081914178453,
    -0.8314696123025452, 0.5555702330196022, -0.8448535652497071, 0.5349976198870973,
    -0.8577286100002721, 0.5141027441932218, -0.8700869911087115, 0.49289819222978404,
    -0.881921264348355, 0.47139673682599764, -0.8932243011955153, 0.4496113296546066,
    -0.9039892931234433, 0.4275550934302821, -0.9142097557035307, 0.40524131400498986,
    -0.9238795325112867, 0.3826834323650898, -0.9329927988347388, 0.35989503653498817,
    -0.9415440651830208, 0.33688985339222005, -0.9495281805930367, 0.31368174039889146,
    -0.9569403357322088, 0.2902846772544624, -0.9637760657954398, 0.26671275747489837,
    -0.970031253194544, 0.2429801799032639, -0.9757021300385286, 0.2191012401568698,
    -0.9807852804032304, 0.19509032201612828, -0.9852776423889412, 0.17096188876030122,
    -0.989176509964781, 0.14673047445536175, -0.99247953459871, 0.1224106751992162,
    -0.9951847266721969, 0.0980171403295606, -0.9972904566786902, 0.07356456359966743,
    -0.9987954562051724, 0.049067674327418015, -0.9996988186962042, 0.024541228522912288,
    -1.0, 4.1340642196527976e-43, -0.9996988186962042, -0.024541228522912288,
    -0.9987954562051724, -0.049067674327418015, -0.9972904566786902, -0.07356456359966743,
    -0.9951847266721969, -0.0980171403295606, -0.99247953459871, -0.1224106751992162,
    -0.989176509964781, -0.14673047445536175, -0.9852776423889412, -0.17096188876030122,
    -0.9807852804032304, -0.19509032201612828, -0.9757021300385286, -0.2191012401568698,
    -0.970031253194544, -0.2429801799032639, -0.9637760657954398, -0.26671275747489837,
    -0.9569403357322088, -0.2902846772544624, -0.9495281805930367, -0.31368174039889146,
    -0.9415440651830208, -0.33688985339222005, -0.9329927988347388, -0.35989503653498817,
    -0.9238795325112867, -0.3826834323650898, -0.9142097557035307, -0.40524131400498986,
    -0.9039892931234433, -0.4275550934302821, -0.8932243011955153, -0.4496113296546066,
    -0.881921264348355, -0.47139673682599764, -0.8700869911087115, -0.49289819222978404,
    -0.8577286100002721, -0.5141027441932218, -0.8448535652497071, -0.5349976198870973,
    -0.8314696123025452, -0.5555702330196022, -0.8175848131515837, -0.5758081914178453,
    -0.8032075314806449, -0.5956993044924334, -0.7883464276266062, -0.6152315905806268,
    -0.773010453362737, -0.6343932841636455, -0.7572088465064846, -0.6531728429537768,
    -0.7409511253549591, -0.6715589548470184, -0.7242470829514669, -0.6895405447370669,
    -0.7071067811865476, -0.7071067811865476, -0.6895405447370669, -0.7242470829514669,
    -0.6715589548470184, -0.7409511253549591, -0.6531728429537768, -0.7572088465064846,
    -0.6343932841636455, -0.773010453362737, -0.6152315905806268, -0.7883464276266062,
    -0.5956993044924334, -0.8032075314806449, -0.5758081914178453, -0.8175848131515837,
    -0.5555702330196022, -0.8314696123025452, -0.5349976198870973, -0.8448535652497071,
    -0.5141027441932218, -0.8577286100002721, -0.49289819222978404, -0.8700869911087115,
    -0.47139673682599764, -0.881921264348355, -0.4496113296546066, -0.8932243011955153,
    -0.4275550934302821, -0.9039892931234433, -0.40524131400498986, -0.9142097557035307,
    -0.3826834323650898, -0.9238795325112867, -0.35989503653498817, -0.9329927988347388,
    -0.33688985339222005, -0.9415440651830208, -0.31368174039889146, -0.9495281805930367,
    -0.2902846772544624, -0.9569403357322088, -0.26671275747489837, -0.9637760657954398,
    -0.2429801799032639, -0.970031253194544, -0.2191012401568698, -0.9757021300385286,
    -0.19509032201612828, -0.9807852804032304, -0.17096188876030122, -0.9852776423889412,
    -0.14673047445536175, -0.989176509964781, -0.1224106751992162, -0.99247953459871,
    -0.0980171403295606, -0.9951847266721969, -0.07356456359966743, -0.9972904566786902,
    -0.049067674327418015, -0.9987954562051724, -0.024541228522912288, -0.9996988186962042,
    2.2338764406549882e-41, -1.0, 0.024541228522912288, -0.9996988186962042,
    0.049067674327418015, -0.9987954562051724, 0.07356456359966743, -0.9972904566786902,
    0.0980171403295606, -0.9951847266721969, 0.1224106751992162, -0.99247953459871,
    0.14673047445536175, -0.989176509964781, 0.17096188876030122, -0.9852776423889412,
    0.19509032201612828, -0.9807852804032304, 0.2191012401568698, -0.9757021300385286,
    0.2429801799032639, -0.970031253194544, 0.26671275747489837, -0.9637760657954398,
    0.2902846772544624, -0.9569403357322088, 0.31368174039889146, -0.9495281805930367,
    0.33688985339222005, -0.9415440651830208, 0.35989503653498817, -0.9329927988347388,
    0.3826834323650898, -0.9238795325112867, 0.40524131400498986, -0.9142097557035307,
    0.4275550934302821, -0.9039892931234433, 0.4496113296546066, -0.8932243011955153,
    0.47139673682599764, -0.881921264348355, 0.49289819222978404, -0.8700869911087115,
    0.5141027441932218, -0.8577286100002721, 0.5349976198870973, -0.8448535652497071,
    0.5555702330196022, -0.8314696123025452, 0.5758081914178453, -0.8175848131515837,
    0.5956993044924334, -0.8032075314806449, 0.6152315905806268, -0.7883464276266062,
    0.6343932841636455, -0.773010453362737, 0.6531728429537768, -0.7572088465064846,
    0.6715589548470184, -0.7409511253549591, 0.6895405447370669, -0.7242470829514669,
    0.7071067811865476, -0.7071067811865476, 0.7242470829514669, -0.6895405447370669,
    0.7409511253549591, -0.6715589548470184, 0.7572088465064846, -0.6531728429537768,
    0.773010453362737, -0.6343932841636455, 0.7883464276266062, -0.6152315905806268,
    0.8032075314806449, -0.5956993044924334, 0.8175848131515837, -0.5758081914178453,
    0.8314696123025452, -0.5555702330196022, 0.8448535652497071, -0.5349976198870973,
    0.8577286100002721, -0.5141027441932218, 0.8700869911087115, -0.49289819222978404,
    0.881921264348355, -0.47139673682599764, 0.8932243011955153, -0.4496113296546066,
    0.9039892931234433, -0.4275550934302821, 0.9142097557035307, -0.40524131400498986,
    0.9238795325112867, -0.3826834323650898, 0.9329927988347388, -0.35989503653498817,
    0.9415440651830208, -0.33688985339222005, 0.9495281805930367, -0.31368174039889146,
    0.9569403357322088, -0.2902846772544624, 0.9637760657954398, -0.26671275747489837,
    0.970031253194544, -0.2429801799032639, 0.9757021300385286, -0.2191012401568698,
    0.9807852804032304, -0.19509032201612828, 0.9852776423889412, -0.17096188876030122,
    0.989176509964781, -0.14673047445536175, 0.99247953459871, -0.1224106751992162,
    0.9951847266721969, -0.0980171403295606, 0.9972904566786902, -0.07356456359966743,
    0.9987954562051724, -0.049067674327418015, 0.9996988186962042, -0.024541228522912288,
};

// x = k h + b: the table entry of k and b; false when x is out of range.
__device__ static inline bool ss_trig_reduce(double x, double* ca, double* sa, double* b)
{
    if (!(fabs(x) < 1e15)) return false;
    const double k = rint(x * SS_TRIG_INV_H);
    *b = fma(-k, SS_TRIG_H_LO, fma(-k, SS_TRIG_H_HI, x));
    const int j = (int)((long long)k & 255);
    *ca = ss_trig_table[2 * j];
    *sa = ss_trig_table[2 * j + 1];
    return true;
}

// cm = cos b - 1 and sm = sin b / b - 1 in float32.
__device__ static inline void ss_trig_small(double b, double* cm, double* sm)
{
    const float bf = (float)b;
    const float b2 = bf * bf;
    *cm = (double)(b2 * fmaf(b2, 1.0f / 24.0f, -0.5f));
    *sm = (double)(b2 * fmaf(b2, 1.0f / 120.0f, -1.0f / 6.0f));
}

__device__ static inline double ss_cos32(double x)
{
    double ca, sa, b, cm, sm;
    if (!ss_trig_reduce(x, &ca, &sa, &b)) return cos(x);
    ss_trig_small(b, &cm, &sm);
    return fma(-sa, fma(b, sm, b), fma(ca, cm, ca));
}

__device__ static inline double ss_sin32(double x)
{
    double ca, sa, b, cm, sm;
    if (!ss_trig_reduce(x, &ca, &sa, &b)) return sin(x);
    ss_trig_small(b, &cm, &sm);
    return fma(ca, fma(b, sm, b), fma(sa, cm, sa));
}

// ---- float64 sin and cos for field callbacks (precisions other than fp32) ---
// The reduction and table of ss_trig_reduce (x = k h + b, |b| <= h / 2 =
// pi / 256), then cos b - 1 and sin b / b - 1 by their Taylor series to b^6
// in float64 (the first terms left out, b^8/8! and b^8/9!, are below 2e-20):
//   cos x = cos a + cos a (cos b - 1) - sin a b (1 + (sin b / b - 1)).
// Arguments beyond the reduction's range (and NaN) take the library functions.
__device__ static inline void ss_trig_small64(double b, double* cm, double* sm)
{
    const double b2 = b * b;
    *cm = b2 * fma(b2, fma(b2, -1.0 / 720.0, 1.0 / 24.0), -0.5);
    *sm = b2 * fma(b2, fma(b2, -1.0 / 5040.0, 1.0 / 120.0), -1.0 / 6.0);
}

__device__ static inline double ss_cos64(double x)
{
    double ca, sa, b, cm, sm;
    if (!ss_trig_reduce(x, &ca, &sa, &b)) return cos(x);
    ss_trig_small64(b, &cm, &sm);
    return fma(-sa, fma(b, sm, b), fma(ca, cm, ca));
}

__device__ static inline double ss_sin64(double x)
{
    double ca, sa, b, cm, sm;
    if (!ss_trig_reduce(x, &ca, &sa, &b)) return sin(x);
    ss_trig_small64(b, &cm, &sm);
    return fma(ca, fma(b, sm, b), fma(sa, cm, sa));
}

// sin and cos of x together: one reduction, one table load.
__device__ static inline void ss_sincos64(double x, double* s, double* c)
{
    double ca, sa, b, cm, sm;
    if (!ss_trig_reduce(x, &ca, &sa, &b)) {
        sincos(x, s, c);
        return;
    }
    ss_trig_small64(b, &cm, &sm);
    const double bs = fma(b, sm, b);
    *s = fma(ca, bs, fma(sa, cm, sa));
    *c = fma(-sa, bs, fma(ca, cm, ca));
}

// 2^(j/64), j = 0 .. 63
__constant__ static const double ss_exp_table[64] = {
    1.0, 1.0108892860517005, 1.0218971486541166, 1.0330248790212284,
    1.0442737824274138, 1.0556451783605572, 1.0671404006768237, 1.0787607977571199,
    1.0905077326652577, 1.102382583307841, 1.1143867425958924, 1.1265216186082418,
    1.1387886347566916, 1.1511892299529827, 1.1637248587775775, 1.1763969916502812,
    1.189207115002721, 1.202156731452703, 1.215247359980469, 1.22848053610687,
    1.241857812073484, 1.255380757024691, 1.2690509571917332, 1.2828700160787783,
    1.2968395546510096, 1.3109612115247644, 1.3252366431597413, 1.339667524053303,
    1.3542555469368927, 1.3690024229745905, 1.383909881963832, 1.3989796725383112,
    1.4142135623730951, 1.42961333839197, 1.4451808069770467, 1.460917794180647,
    1.4768261459394993, 1.4929077282912648, 1.5091644275934228, 1.5255981507445384,
    1.5422108254079407, 1.559004400237837, 1.5759808451078865, 1.593142151342267,
    1.6104903319492543, 1.6280274218573478, 1.645755478153965, 1.6636765803267364,
    1.681792830507429, 1.7001063537185235, 1.718619298122478, 1.7373338352737062,
    1.7562521603732995, 1.7753764925265212, 1.7947090750031072, 1.8142521755003989,
    1.8340080864093424, 1.8539791250833855, 1.8741676341103, 1.8945759815869656,
    1.9152065613971474, 1.9360617934922943, 1.9571441241754002, 1.978456026387951,
};

// exp x = 2^n 2^(j/64) exp(b): x = (64 n + j) ln2/64 + b, |b| <= ln2/128, and
// exp(b) = 1 + b + b^2 q with q = 1/2 + b/6 + b^2/24 + b^3/120 in float32
// (b^2 q < 1.5e-5: its rounding is about 1e-12 relative in the result).
// Below -746 the float64 exponential is 0 exactly; beyond [-708, 709] (and
// NaN) the float64 function is used.
__device__ static inline double ss_exp32(double x)
{
    if (x < -746.0) return 0.0;
    if (!(x >= -708.0 && x <= 709.0)) return exp(x);
    const double m = rint(x * SS_EXP_INV_H);
    const double b = fma(-m, SS_EXP_H_LO, fma(-m, SS_EXP_H_HI, x));
    const int mi = (int)m;
    const double t = ss_exp_table[mi & 63];
    const float bf = (float)b;
    const float q = fmaf(bf, fmaf(bf, fmaf(bf, 1.0f / 120.0f, 1.0f / 24.0f), 1.0f / 6.0f), 0.5f);
    const double em = b + (double)(bf * bf * q);
    return ldexp(fma(t, em, t), mi >> 6);
}

#define SS_CONTEXT const ss_ctx* ss_ctx_ref
#define SS_SAMPLE(k, t) ss_interp_sample(ss_ctx_ref, (k), (t))
#define SS_SAMPLE_AXES(k, t, out, n) ss_interp_sample_axes(ss_ctx_ref, (k), (t), (out), (n))
#define SS_SAMPLE_AXIS(k, a, t) ss_interp_sample_axis(ss_ctx_ref, (k), (a), (t))

// find_time_index, spinsim/__init__.py:2026-2061: uniform-grid guess, checked,
// then bisection capped at `size` iterations; -1 when out of range.
__device__ static inline long long ss_find_time_index(const ss_table* tb, double t, double* param)
{
    const double* time = tb->time;
    const long long size = tb->size;
    if (time[0] <= t && t <= time[size - 1]) {
        // A guess outside [0, size) (a repeated first knot gives inf) goes
        // straight to the bisection instead of an undefined conversion.
        const double guess = (t - time[0]) / (time[1] - time[0]);
        long long i = (guess >= 0.0 && guess < (double)size) ? (long long)guess : size;
        bool low = false, high = true;
        if (i + 1 < size) {
            low = time[i] <= t;
            high = t <= time[i + 1];
        } else {
            i = size;
        }
        if (!(low && high)) {
            long long imax = size - 1, imin = 0, budget = size;
            while (!(low && high)) {
                if (!low) imax = i; else imin = i;
                i = (long long)((double)(imax + imin) / 2);
                low = false;
                high = true;
                if (i + 1 < size) {
                    low = time[i] <= t;
                    high = t <= time[i + 1];
                }
                if (--budget == 0) { *param = 0.0; return -1; }
            }
        }
        *param = (t - time[i]) / (time[i + 1] - time[i]);
        return i;
    }
    *param = 0.0;
    return -1;
}

// Cubic Hermite tangents of interval i scaled by its width (m_left, m_right of
// :2107-2136: averaged neighbouring secants, the secant itself at the ends,
// zero next to a repeated knot).
__device__ static inline void ss_interp_tangents(const double* time, const double* amp, long long size,
                                                 long long i, double* left, double* right)
{
    const double h = time[i + 1] - time[i];
    const double mid = (time[i] != time[i + 1]) ? (amp[i + 1] - amp[i]) / h : 0.0;
    double l, r;
    if (i == 0) l = mid;
    else if (time[i - 1] != time[i])
        l = 0.5 * ((amp[i] - amp[i - 1]) / (time[i] - time[i - 1]) + mid);
    else l = 0.0;
    if (i + 2 == size) r = mid;
    else if (time[i + 1] != time[i + 2])
        r = 0.5 * ((amp[i + 2] - amp[i + 1]) / (time[i + 2] - time[i + 1]) + mid);
    else r = 0.0;
    *left = l * h;
    *right = r * h;
}

// sampler(), spinsim/__init__.py:2065-2142 (orders 0, 1 and 3).  Linear and
// cubic intervals are remembered per lane (ss_ctx::line) and re-evaluated as
// polynomials in p while t stays inside them; the values agree with the
// reference's form to rounding (~1e-16 relative).
__device__ static inline double ss_interp_sample(const ss_ctx* ctx, int k, double t)
{
    if (k < SS_INTERP_LINES) {
        const ss_interp_line& ln = ctx->line[k];
        const double p = (t - ln.t0) * ln.inv_h;
        if (p >= 0.0 && p <= 1.0) return ((ln.c3 * p + ln.c2) * p + ln.c1) * p + ln.c0;
    }
    const ss_table* tb = &ctx->tables[k];
    const double* time = tb->time;
    const double* amp = tb->amplitude + (long long)ctx->sim * tb->amplitude_stride;
    const long long size = tb->size;
    if (size < 2) return (size == 1 && !tb->zero_boundary) ? amp[0] : 0.0;  // degenerate table
    if (!tb->zero_boundary) {
        if (t <= time[0]) return amp[0];
        if (t >= time[size - 1]) return amp[size - 1];
    }
    double p;
    const long long i = ss_find_time_index(tb, t, &p);
    double out = 0.0;
    if (tb->order == 0) {
        if (i >= 0) out = amp[i];
    } else if (tb->order == 1) {
        if (i >= 0 && i + 1 < size) {
            out = amp[i] * (1 - p) + amp[i + 1] * p;
            const double h = time[i + 1] - time[i];
            if (k < SS_INTERP_LINES && h > 0.0) {
                ss_interp_line& ln = ctx->line[k];
                ln.t0 = time[i];
                ln.inv_h = 1.0 / h;
                ln.idx = i;
                ln.c0 = amp[i];
                ln.c1 = amp[i + 1] - amp[i];
                ln.c2 = 0.0;
                ln.c3 = 0.0;
            }
        }
    } else if (i >= 0 && i + 1 < size) {
        const double h = time[i + 1] - time[i];
        double left, right;
        ss_interp_tangents(time, amp, size, i, &left, &right);
        const double p2 = p * p, p3 = p2 * p;
        out = amp[i] * (2 * p3 - 3 * p2 + 1) + amp[i + 1] * (-2 * p3 + 3 * p2)
            + left * (p3 - 2 * p2 + p) + right * (p3 - p2);
        if (k < SS_INTERP_LINES && h > 0.0) {
            ss_interp_line& ln = ctx->line[k];
            ln.t0 = time[i];
            ln.inv_h = 1.0 / h;
            ln.idx = i;
            ln.c0 = amp[i];
            ln.c1 = left;
            ln.c2 = 3.0 * (amp[i + 1] - amp[i]) - 2.0 * left - right;
            ln.c3 = 2.0 * (amp[i] - amp[i + 1]) + left + right;
        }
    }
    return out;
}

// One amplitude row of table tb at interval i, parameter p (i < 0: outside
// the table), :2065-2142.
__device__ static inline double ss_interp_row(const ss_table* tb, const double* amp, long long i, double p)
{
    const long long size = tb->size;
    if (i < 0) return 0.0;
    if (tb->order == 0) return amp[i];
    if (i + 1 >= size) return 0.0;
    if (tb->order == 1) return amp[i] * (1 - p) + amp[i + 1] * p;
    double left, right;
    ss_interp_tangents(tb->time, amp, size, i, &left, &right);
    const double p2 = p * p, p3 = p2 * p;
    return amp[i] * (2 * p3 - 3 * p2 + 1) + amp[i + 1] * (-2 * p3 + 3 * p2) + left * (p3 - 2 * p2 + p)
           + right * (p3 - p2);
}

// Components [a0, a1) of multi-axis table k at time t into out[0 .. a1-a0),
// from one index search; component a equals the one-axis sampler of amplitude
// row a.  (The reference rejects 2-D amplitudes and suggests one sampler per
// axis, spinsim/__init__.py:2062-2063; here the axes share the search.)
__device__ static inline void ss_interp_sample_range(const ss_ctx* ctx, int k, double t, int a0, int a1,
                                                     double* out)
{
    const ss_table* tb = &ctx->tables[k];
    if (a1 > tb->axes) a1 = tb->axes;
    const double* time = tb->time;
    const double* amp0 = tb->amplitude + (long long)ctx->sim * tb->amplitude_stride;
    const long long size = tb->size;
    const long long last = size - 1;
    if (size < 2) {
        for (int a = a0; a < a1; ++a)
            out[a - a0] = (size == 1 && !tb->zero_boundary) ? amp0[a * tb->axis_stride] : 0.0;
        return;
    }
    if (!tb->zero_boundary && (t <= time[0] || t >= time[last])) {
        const long long j = t <= time[0] ? 0 : last;
        for (int a = a0; a < a1; ++a) out[a - a0] = amp0[a * tb->axis_stride + j];
        return;
    }
    double p;
    const long long i = ss_find_time_index(tb, t, &p);
    for (int a = a0; a < a1; ++a) out[a - a0] = ss_interp_row(tb, amp0 + a * tb->axis_stride, i, p);
}

// The first n components of multi-axis table k at time t (n at most the
// table's axes; further entries of out are left alone).
__device__ static inline void ss_interp_sample_axes(const ss_ctx* ctx, int k, double t, double* out, int n)
{
    ss_interp_sample_range(ctx, k, t, 0, n, out);
}

// Component a of multi-axis table k at time t (0 beyond the table's axes).
__device__ static inline double ss_interp_sample_axis(const ss_ctx* ctx, int k, int a, double t)
{
    double v = 0.0;
    if (a >= 0) ss_interp_sample_range(ctx, k, t, a, a + 1, &v);
    return v;
}
