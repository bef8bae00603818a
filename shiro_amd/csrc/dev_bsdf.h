// PBR BSDF of Shiro's CPU path for the kernels: the layered
// UsdPreviewSurface / standard_surface / OpenPBR-parameter model of the
// reference (src/backend/cpu/CpuPathTracer.cpp:841-1441): Oren-Nayar diffuse
// with wrap, GGX base and coat lobes with Schlick Fresnel and height-
// correlated Smith G2, thin film tint, sheen, rough dielectric transmission.
// Material-only quantities (lobe weights, F0s, tints) come precomputed from
// the host (MaterialDerived) with the same expressions; everything that
// depends on directions is evaluated here in the reference's order.
#pragma once

#include "dev_math.h"
#include "layout.h"

namespace shiro_hip {

// A material at a hit: its parameters, the derived constants shading
// evaluates (d) and those sampling decisions and pdfs read (ds).  d and ds
// are the same table entry except at hits on textured materials whose
// sampling texels were frozen (layout.h DevTexture).
struct Mat {
    const shiro_hip_material* m;
    const MaterialDerived* d;
    const MaterialDerived* ds;
};

DEV V3 tangent_of(V3 n) {
    const V3 axis = fabsf(n.z) < 0.999f ? V3{0.0f, 0.0f, 1.0f} : V3{0.0f, 1.0f, 0.0f};
    return normalize(cross(axis, n));
}

DEV V3 projected_tangent(const shiro_hip_material& m, V3 n) {
    if (m.hasTangentOverride) {
        const V3 t0 = ld3(m.tangentOverride);
        const V3 t = t0 - n * dot(t0, n);
        if (len(t) > 0.0f) return normalize(t);
    }
    return tangent_of(n);
}

// CpuPathTracer.cpp:862-879.
DEV float anisotropy_phase(V3 n, V3 t, V3 v, V3 l, float rot) {
    const V3 td = safe_normalize(t - n * dot(t, n), tangent_of(n));
    const V3 bt = normalize(cross(n, td));
    const V3 h = safe_normalize(v + l, n);
    const V3 hp = h - n * dot(n, h);
    if (len(hp) <= 1.0e-6f) return cosf(rot * 2.0f * kPi);
    const V3 hd = normalize(hp);
    const float ang = atan2f(dot(hd, bt), dot(hd, td));
    return cosf(ang - rot * 2.0f * kPi);
}

// CpuPathTracer.cpp:881-893.  With zero anisotropy the phase is multiplied
// by zero and the result is clamp(r, 0.02, 1) exactly, so it is skipped.
// PLAIN: the scene's materials have no specular or coat anisotropy, no thin
// film, no sheen and no subsurface wrap (DevScene::plainBsdf), so those
// terms are known and their code is left out; the results are those of the
// general evaluation for such materials.
template <bool PLAIN = false>
DEV float roughness_eval(float r, float aniso, float rot, const shiro_hip_material& m, V3 n, V3 v, V3 l) {
    const float a = clampf(aniso, -0.95f, 0.95f);
    if (PLAIN || a == 0.0f) return clampf(r * (1.0f - 0.35f * a * 1.0f), 0.02f, 1.0f);
    const float phase = anisotropy_phase(n, projected_tangent(m, n), v, l, rot);
    return clampf(r * (1.0f - 0.35f * a * phase), 0.02f, 1.0f);
}

DEV float ggx_d(float nDotH, float roughness) {
    const float a = fmaxs(0.02f, roughness * roughness);
    const float a2 = a * a;
    const float den = fmaxs(1.0e-6f, sq(nDotH * nDotH * (a2 - 1.0f) + 1.0f));
    return a2 / (kPi * den);
}

DEV float smith_lambda(float c, float a2) {
    const float cc = fmaxs(c, 1.0e-6f);
    return (-1.0f + sqrtf(1.0f + a2 * (1.0f - cc * cc) / (cc * cc))) * 0.5f;
}

DEV float ggx_g2(float nDotV, float nDotL, float roughness) {
    const float a = fmaxs(0.02f, roughness * roughness);
    const float a2 = a * a;
    return 1.0f / (1.0f + smith_lambda(nDotV, a2) + smith_lambda(nDotL, a2));
}

DEV V3 schlick(V3 f0, float c) {
    const float x = clampf(1.0f - c, 0.0f, 1.0f);
    const float x5 = x * x * x * x * x;
    return f0 + (V3{1.0f, 1.0f, 1.0f} - f0) * x5;
}

// Exact unpolarised dielectric Fresnel (CpuPathTracer.cpp:1011-1033).
DEV float fresnel_dielectric(float cosI, float etaI, float etaT) {
    float c = clampf(cosI, -1.0f, 1.0f);
    float ei = fmaxs(etaI, 1.0e-6f);
    float et = fmaxs(etaT, 1.0e-6f);
    if (c <= 0.0f) {
        const float tmp = ei;
        ei = et;
        et = tmp;
        c = fabsf(c);
    }
    const float si = sqrtf(fmaxs(0.0f, 1.0f - c * c));
    const float stt = ei / et * si;
    if (stt >= 1.0f) return 1.0f;
    const float ct = sqrtf(fmaxs(0.0f, 1.0f - stt * stt));
    const float rpa = ((et * c) - (ei * ct)) / fmaxs((et * c) + (ei * ct), 1.0e-6f);
    const float rpe = ((ei * c) - (et * ct)) / fmaxs((ei * c) + (et * ct), 1.0e-6f);
    return 0.5f * (rpa * rpa + rpe * rpe);
}

DEV V3 one_minus_clamped(V3 v) {
    return V3{clampf(1.0f - v.x, 0.0f, 1.0f), clampf(1.0f - v.y, 0.0f, 1.0f), clampf(1.0f - v.z, 0.0f, 1.0f)};
}

DEV V3 layer_transmittance(V3 f0, float nDotV, float nDotL) {
    const V3 fv = schlick(f0, clampf(nDotV, 0.0f, 1.0f));
    const V3 fl = schlick(f0, clampf(nDotL, 0.0f, 1.0f));
    return one_minus_clamped((fv + fl) * 0.5f);
}

DEV V3 interface_transmittance(V3 f0, float ca, float cb) {
    const V3 fa = schlick(f0, clampf(fabsf(ca), 0.0f, 1.0f));
    const V3 fb = schlick(f0, clampf(fabsf(cb), 0.0f, 1.0f));
    return one_minus_clamped((fa + fb) * 0.5f);
}

DEV V3 coat_transmission_scale(const Mat& M, V3 cn, V3 v, V3 l) {
    return ld3(M.d->coatTint) * interface_transmittance(ld3(M.d->coatF0), dot(cn, v), dot(cn, l));
}

// Microfacet reflection (CpuPathTracer.cpp:1096-1115).
DEV V3 specular_lobe(V3 n, V3 v, V3 l, V3 f0, float r) {
    const float nv = clampf(dot(n, v), 0.0f, 1.0f);
    const float nl = clampf(dot(n, l), 0.0f, 1.0f);
    if (nv <= 0.0f || nl <= 0.0f) return V3{0.0f, 0.0f, 0.0f};
    const V3 h = safe_normalize(v + l, n);
    const float nh = clampf(dot(n, h), 0.0f, 1.0f);
    const float vh = clampf(dot(v, h), 0.0f, 1.0f);
    const float D = ggx_d(nh, r);
    const float G = ggx_g2(nv, nl, r);
    return schlick(f0, vh) * (D * G / fmaxs(4.0f * nv * nl, 1.0e-6f));
}

DEV float specular_pdf(V3 n, V3 v, V3 l, float r) {
    const float nv = clampf(dot(n, v), 0.0f, 1.0f);
    const float nl = clampf(dot(n, l), 0.0f, 1.0f);
    if (nv <= 0.0f || nl <= 0.0f) return 0.0f;
    const V3 h = safe_normalize(v + l, n);
    const float nh = clampf(dot(n, h), 0.0f, 1.0f);
    const float vh = clampf(dot(v, h), 1.0e-6f, 1.0f);
    return ggx_d(nh, r) * nh / fmaxs(4.0f * vh, 1.0e-6f);
}

// CpuPathTracer.cpp:934-968.
template <bool PLAIN = false>
DEV V3 apply_thin_film(V3 lobe, const shiro_hip_material& m, V3 n, V3 v, V3 l) {
    if (PLAIN || m.thinFilmThickness <= 0.0f) return lobe;
    const V3 h = safe_normalize(v + l, n);
    const float vDotH = clampf(dot(v, h), 0.0f, 1.0f);
    const float thickness = m.thinFilmThickness;
    const float i = fmaxs(m.thinFilmIor, 1.0f);
    const float path = 2.0f * i * thickness * clampf(vDotH, 0.05f, 1.0f);
    const float b0 = 0.5f + 0.5f * cosf(4.0f * kPi * path / fmaxs(650.0f, 1.0f));
    const float b1 = 0.5f + 0.5f * cosf(4.0f * kPi * path / fmaxs(510.0f, 1.0f));
    const float b2 = 0.5f + 0.5f * cosf(4.0f * kPi * path / fmaxs(475.0f, 1.0f));
    const float strength = clampf(thickness / 1200.0f, 0.0f, 1.0f);
    const V3 tint = lerp(V3{1.0f, 1.0f, 1.0f}, (V3{1.0f, 1.0f, 1.0f} + V3{b0, b1, b2}) * 0.5f, strength);
    return lobe * tint;
}

// Reflective mixture pdf (CpuPathTracer.cpp:1232-1273).  The out-of-line
// BSDF functions take the material as separate pointers, which travel in
// registers (a Mat passed by value goes through the stack).
template <bool PLAIN>
DEV_CALL float bsdf_pdf_at(const shiro_hip_material* mp, const MaterialDerived* dsp, V3 n, V3 cn, V3 v, V3 l) {
    const float nl = clampf(dot(n, l), 0.0f, 1.0f);
    if (nl <= 0.0f) return 0.0f;
    const shiro_hip_material& m = *mp;
    const MaterialDerived& ds = *dsp;
    const float wd = ds.wDiffuse, ws = ds.wBaseSpecular, wc = ds.wCoat;
    const float wr = wd + ws + wc;
    if (wr <= 0.0f) return 0.0f;
    const float pd = wd / wr, ps = ws / wr, pc = wc / wr;
    const float rs = roughness_eval<PLAIN>(ds.baseRoughness, m.specularAnisotropy, m.specularRotation, m, n, v, l);
    float coatTerm = 0.0f;
    if (pc != 0.0f) {
        const float rc = roughness_eval<PLAIN>(m.coatRoughness, m.coatAnisotropy, m.coatRotation, m, cn, v, l);
        coatTerm = pc * specular_pdf(cn, v, l, rc);
    }
    return pd * nl * kInvPi + ps * specular_pdf(n, v, l, rs) + coatTerm;
}

// Layered BSDF value (CpuPathTracer.cpp:1275-1354).
template <bool PLAIN>
DEV_CALL V3 bsdf_eval_at(const shiro_hip_material* mp, const MaterialDerived* dp, V3 n, V3 cn, V3 v, V3 l) {
    const float nv = clampf(dot(n, v), 0.0f, 1.0f);
    const float nl = clampf(dot(n, l), 0.0f, 1.0f);
    if (nv <= 0.0f || nl <= 0.0f) return V3{0.0f, 0.0f, 0.0f};
    const shiro_hip_material& m = *mp;
    const MaterialDerived& d = *dp;
    const V3 f0b = ld3(d.specularF0);
    const V3 f0c = ld3(d.coatF0);
    const V3 underCoat = ld3(d.coatTint) * layer_transmittance(f0c, dot(cn, v), dot(cn, l));
    const V3 underBase = layer_transmittance(f0b, nv, nl);
    const V3 vp = normalize(v - n * nv);
    const V3 lp = normalize(l - n * nl);
    const float cosPhi = len(vp) > 0.0f && len(lp) > 0.0f ? clampf(dot(vp, lp), -1.0f, 1.0f) : 0.0f;
    const float sg = clampf(m.diffuseRoughness, 0.0f, 1.0f);
    const float sg2 = sg * sg;
    const float A = 1.0f - (sg2 / (2.0f * (sg2 + 0.33f)));
    const float B = 0.45f * sg2 / (sg2 + 0.09f);
    const float sv = sqrtf(fmaxs(0.0f, 1.0f - nv * nv));
    const float sl = sqrtf(fmaxs(0.0f, 1.0f - nl * nl));
    const float sinAlpha = fmaxs(sv, sl);
    const float tanBeta = fmins(sv / fmaxs(nv, 1.0e-6f), sl / fmaxs(nl, 1.0e-6f));
    // With no subsurface wrap (0 < nl <= 1 here), wnl = clamp((nl + 0) / 1) = nl.
    float wnl = nl;
    if (!PLAIN) {
        const float wrap = clampf(m.subsurface * m.subsurfaceScale *
                                      (0.35f + 0.15f * clampf(m.subsurfaceAnisotropy, -1.0f, 1.0f) +
                                       0.1f * clampf(max3(ld3(m.subsurfaceRadius)) / 3.0f, 0.0f, 1.0f)),
                                  0.0f, 0.9f);
        wnl = clampf((nl + wrap) / (1.0f + wrap), 0.0f, 1.0f);
    }
    const float dterm = (A + B * fmaxs(0.0f, cosPhi) * sinAlpha * tanBeta) * (wnl / fmaxs(nl, 1.0e-6f));
    const V3 diffuse = ld3(d.diffuseColor) * underCoat * underBase * (kInvPi * dterm);
    const float rs = roughness_eval<PLAIN>(d.baseRoughness, m.specularAnisotropy, m.specularRotation, m, n, v, l);
    const V3 spec = underCoat * apply_thin_film<PLAIN>(specular_lobe(n, v, l, f0b, rs), m, n, v, l);
    const float rc = roughness_eval<PLAIN>(m.coatRoughness, m.coatAnisotropy, m.coatRotation, m, cn, v, l);
    const V3 coat = apply_thin_film<PLAIN>(specular_lobe(cn, v, l, f0c, rc), m, cn, v, l);
    V3 out = diffuse + spec + coat;
    if (!PLAIN && (d.flags & kMatSheen)) {
        const V3 h = safe_normalize(v + l, n);
        const float sheenW = clampf(m.sheen, 0.0f, 1.0f);
        const float lh = clampf(dot(l, h), 0.0f, 1.0f);
        const float sheenF = powf(1.0f - lh, 5.0f) * (1.0f - 0.5f * clampf(m.sheenRoughness, 0.0f, 1.0f));
        out = out + ld3(m.sheenColor) * sheenW * sheenF * underCoat * underBase;
    }
    return out;
}

template <bool PLAIN = false>
DEV float bsdf_pdf(const Mat& M, V3 n, V3 cn, V3 v, V3 l) {
    return bsdf_pdf_at<PLAIN>(M.m, M.ds, n, cn, v, l);
}
template <bool PLAIN = false>
DEV V3 bsdf_eval(const Mat& M, V3 n, V3 cn, V3 v, V3 l) {
    return bsdf_eval_at<PLAIN>(M.m, M.d, n, cn, v, l);
}

// Derivatives of bsdf_eval at fixed directions with respect to the base
// colour (per channel: d f_c / d baseColor_c) and the roughness
// (d f_c / d roughness), for the differentiable class of the texture
// extension: no specular anisotropy, thin film, sheen or transmission
// (shiro_hip_render_adjoint rejects other textured materials).  The base
// colour enters through the diffuse colour and, for metals, the base F0 (in
// the layer transmittance and the Fresnel term); the roughness through the
// GGX D and the height-correlated G2.  Clamps contribute no derivative
// where they are active.
struct BsdfGrad {
    V3 dBase, dRough;
};

DEV float schlick_weight(float c) {
    const float x = clampf(1.0f - c, 0.0f, 1.0f);
    return x * x * x * x * x;
}

// d Lambda(c) / d a2 of smith_lambda.
DEV float smith_lambda_da2(float c, float a2) {
    const float cc = fmaxs(c, 1.0e-6f);
    const float t = (1.0f - cc * cc) / (cc * cc);
    return t / (4.0f * sqrtf(1.0f + a2 * t));
}

DEV BsdfGrad bsdf_grad(const Mat& M, V3 n, V3 cn, V3 v, V3 l) {
    BsdfGrad g{V3{0.0f, 0.0f, 0.0f}, V3{0.0f, 0.0f, 0.0f}};
    const float nv = clampf(dot(n, v), 0.0f, 1.0f);
    const float nl = clampf(dot(n, l), 0.0f, 1.0f);
    if (nv <= 0.0f || nl <= 0.0f) return g;
    const shiro_hip_material& m = *M.m;
    const MaterialDerived& d = *M.d;
    const V3 f0b = ld3(d.specularF0);
    const V3 uC = ld3(d.coatTint) * layer_transmittance(ld3(d.coatF0), dot(cn, v), dot(cn, l));
    // Layer transmittance of the base and its derivative in F0.
    const float wv = schlick_weight(nv), wl = schlick_weight(nl);
    const V3 Fv = schlick(f0b, nv), Fl = schlick(f0b, nl);
    const V3 sumF = (Fv + Fl) * 0.5f;
    const V3 uB = one_minus_clamped(sumF);
    const float dUB = -0.5f * ((1.0f - wv) + (1.0f - wl));
    const V3 dUBdF0{sumF.x > 0.0f && sumF.x < 1.0f ? dUB : 0.0f, sumF.y > 0.0f && sumF.y < 1.0f ? dUB : 0.0f,
                    sumF.z > 0.0f && sumF.z < 1.0f ? dUB : 0.0f};
    // Diffuse factor (dterm / pi) as in bsdf_eval.
    const V3 vp = normalize(v - n * nv);
    const V3 lp = normalize(l - n * nl);
    const float cosPhi = len(vp) > 0.0f && len(lp) > 0.0f ? clampf(dot(vp, lp), -1.0f, 1.0f) : 0.0f;
    const float sg = clampf(m.diffuseRoughness, 0.0f, 1.0f);
    const float sg2 = sg * sg;
    const float A = 1.0f - (sg2 / (2.0f * (sg2 + 0.33f)));
    const float B = 0.45f * sg2 / (sg2 + 0.09f);
    const float sv = sqrtf(fmaxs(0.0f, 1.0f - nv * nv));
    const float sl = sqrtf(fmaxs(0.0f, 1.0f - nl * nl));
    const float tanBeta = fmins(sv / fmaxs(nv, 1.0e-6f), sl / fmaxs(nl, 1.0e-6f));
    const float wrap = clampf(m.subsurface * m.subsurfaceScale *
                                  (0.35f + 0.15f * clampf(m.subsurfaceAnisotropy, -1.0f, 1.0f) +
                                   0.1f * clampf(max3(ld3(m.subsurfaceRadius)) / 3.0f, 0.0f, 1.0f)),
                              0.0f, 0.9f);
    const float wnl = clampf((nl + wrap) / (1.0f + wrap), 0.0f, 1.0f);
    const float kd = kInvPi * (A + B * fmaxs(0.0f, cosPhi) * fmaxs(sv, sl) * tanBeta) * (wnl / fmaxs(nl, 1.0e-6f));
    const float metal = clampf(m.metallic, 0.0f, 1.0f);
    const float trans = clampf(m.transmission, 0.0f, 1.0f);
    const float sss = clampf(m.subsurface, 0.0f, 1.0f);
    const float blend = clampf(max3(ld3(m.subsurfaceRadius)) / 3.0f, 0.0f, 1.0f);
    const float kDiff = ((1.0f - sss) + sss * (1.0f - blend)) *
                        (clampf(m.baseWeight, 0.0f, 1.0f) * (1.0f - metal) * (1.0f - trans));
    const V3 dc = ld3(d.diffuseColor);
    // Specular lobe pieces.
    const V3 h = safe_normalize(v + l, n);
    const float nh = clampf(dot(n, h), 0.0f, 1.0f);
    const float vh = clampf(dot(v, h), 0.0f, 1.0f);
    const float bR = d.baseRoughness;
    const float rs = clampf(bR, 0.02f, 1.0f);
    const float a = fmaxs(0.02f, rs * rs);
    const float a2 = a * a;
    const float q = nh * nh * (a2 - 1.0f) + 1.0f;
    const float den = fmaxs(1.0e-6f, q * q);
    const float D = a2 / (kPi * den);
    const float G = ggx_g2(nv, nl, rs);
    const float den4 = fmaxs(4.0f * nv * nl, 1.0e-6f);
    const float wh = schlick_weight(vh);
    const V3 F = schlick(f0b, vh);
    // d/d base.
    const float dSpecdF0 = (1.0f - wh) * D * G / den4;
    g.dBase = V3{kDiff * uC.x * uB.x * kd + dc.x * uC.x * dUBdF0.x * metal * kd + uC.x * dSpecdF0 * metal,
                 kDiff * uC.y * uB.y * kd + dc.y * uC.y * dUBdF0.y * metal * kd + uC.y * dSpecdF0 * metal,
                 kDiff * uC.z * uB.z * kd + dc.z * uC.z * dUBdF0.z * metal * kd + uC.z * dSpecdF0 * metal};
    // d/d roughness: baseRoughness = clamp(r + coat term, 0, 1), rs = clamp(bR, 0.02, 1),
    // a = max(0.02, rs^2), a2 = a^2.
    if (bR > 0.02f && bR < 1.0f && rs * rs > 0.02f) {
        const float dDda2 = q * q > 1.0e-6f ? (1.0f / (kPi * q * q)) * (1.0f - 2.0f * a2 * nh * nh / q)
                                             : 1.0f / (kPi * 1.0e-6f);
        const float dGda2 = -G * G * (smith_lambda_da2(nv, a2) + smith_lambda_da2(nl, a2));
        const float da2drs = 2.0f * a * 2.0f * rs;
        const float dS = (dDda2 * G + D * dGda2) * da2drs / den4;
        g.dRough = uC * F * dS;
    }
    return g;
}

// Rough dielectric transmission (CpuPathTracer.cpp:1136-1182).
DEV_CALL V3 transmission_lobe_at(const shiro_hip_material* mp, const MaterialDerived* dp, V3 n, V3 cn, V3 v, V3 l,
                                 float etaI, float etaT, float r) {
    const Mat M{mp, dp, dp};
    const float cwo = dot(n, v);
    const float cwi = dot(n, l);
    const float acwo = fabsf(cwo), acwi = fabsf(cwi);
    if (acwo <= 0.0f || acwi <= 0.0f || cwo * cwi >= 0.0f) return V3{0.0f, 0.0f, 0.0f};
    const float eta = etaT / fmaxs(etaI, 1.0e-6f);
    V3 h = safe_normalize(v + l * eta, n);
    if (dot(n, h) < 0.0f) h = h * -1.0f;
    const float oh = dot(v, h);
    const float ih = dot(l, h);
    if (oh * ih >= 0.0f) return V3{0.0f, 0.0f, 0.0f};
    const float nh = clampf(dot(n, h), 0.0f, 1.0f);
    const float sden = oh + eta * ih;
    if (nh <= 0.0f || fabsf(sden) <= 1.0e-6f) return V3{0.0f, 0.0f, 0.0f};
    const float D = ggx_d(nh, r);
    const float G = ggx_g2(acwo, acwi, r);
    const float F = 1.0f - fresnel_dielectric(fabsf(oh), etaI, etaT);
    const shiro_hip_material& m = *M.m;
    const V3 scale = ld3(M.d->transmissionTint) * coat_transmission_scale(M, cn, v, l) *
                     (clampf(m.transmission, 0.0f, 1.0f) * (1.0f - clampf(m.metallic, 0.0f, 1.0f)));
    const float factor = fabsf(D * G * fabsf(ih) * fabsf(oh) / fmaxs(acwo * acwi * sq(sden), 1.0e-6f));
    const float transport = 1.0f / fmaxs(eta * eta, 1.0e-6f);
    return scale * F * (factor * transport);
}

DEV V3 transmission_lobe(const Mat& M, V3 n, V3 cn, V3 v, V3 l, float etaI, float etaT, float r) {
    return transmission_lobe_at(M.m, M.d, n, cn, v, l, etaI, etaT, r);
}

DEV float transmission_pdf(V3 n, V3 v, V3 l, float eta, float r) {
    const float cwo = dot(n, v);
    const float cwi = dot(n, l);
    if (fabsf(cwo) <= 0.0f || fabsf(cwi) <= 0.0f || cwo * cwi >= 0.0f) return 0.0f;
    V3 h = safe_normalize(v + l * eta, n);
    if (dot(n, h) < 0.0f) h = h * -1.0f;
    const float oh = dot(v, h);
    const float ih = dot(l, h);
    const float sden = oh + eta * ih;
    if (oh * ih >= 0.0f || fabsf(sden) <= 1.0e-6f) return 0.0f;
    const float nh = clampf(dot(n, h), 0.0f, 1.0f);
    return ggx_d(nh, r) * nh * fabsf((eta * eta * ih) / sq(sden));
}

DEV bool refract_dir(V3 i, V3 n, float eta, V3* out) {
    const float c = clampf(-dot(i, n), -1.0f, 1.0f);
    const float s2 = fmaxs(0.0f, 1.0f - c * c);
    const float k = 1.0f - eta * eta * s2;
    if (k <= 0.0f) return false;
    *out = normalize(i * eta + n * (eta * c - sqrtf(k)));
    return true;
}

DEV float pane_transmittance(float R) {
    const float r = clampf(R, 0.0f, 1.0f);
    return clampf((1.0f - r) / fmaxs(1.0f + r, 1.0e-6f), 0.0f, 1.0f);
}

DEV V3 sample_cosine(V3 n, V2 u) {
    const float phi = 2.0f * kPi * u.x;
    const float r = sqrtf(u.y);
    const float lx = r * cosf(phi);
    const float ly = r * sinf(phi);
    const float lz = sqrtf(fmaxs(0.0f, 1.0f - u.y));
    const V3 t = tangent_of(n);
    const V3 b = cross(n, t);
    return normalize(t * lx + b * ly + n * lz);
}

DEV V3 sample_ggx_half(V3 n, V2 u, float roughness) {
    const float a = fmaxs(0.02f, roughness * roughness);
    const float a2 = a * a;
    const float phi = 2.0f * kPi * u.x;
    const float ct = sqrtf(fmaxs(0.0f, (1.0f - u.y) / fmaxs(1.0e-6f, 1.0f + (a2 - 1.0f) * u.y)));
    const float st = sqrtf(fmaxs(0.0f, 1.0f - ct * ct));
    const float lx = st * cosf(phi);
    const float ly = st * sinf(phi);
    const V3 t = tangent_of(n);
    const V3 b = cross(n, t);
    return normalize(t * lx + b * ly + n * ct);
}

DEV float power_heuristic(float a, float b) {
    const float a2 = a * a, b2 = b * b;
    return a2 / fmaxs(a2 + b2, 1.0e-6f);
}

// CpuPathTracer.cpp:970-986.
DEV V3 resolve_normal(const shiro_hip_material& m, V3 fallback, bool coat) {
    const bool has = coat ? m.hasCoatNormalOverride : m.hasNormalOverride;
    if (!has) return fallback;
    const V3 c = coat ? ld3(m.coatNormalOverride) : ld3(m.normalOverride);
    if (len(c) <= 0.0f) return fallback;
    V3 n = normalize(c);
    if (dot(n, fallback) < 0.0f) n = n * -1.0f;
    return n;
}

}  // namespace shiro_hip
