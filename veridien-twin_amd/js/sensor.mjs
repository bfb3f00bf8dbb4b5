// Sensor and ISP parameters for the MI355X renderer. The parameter names, defaults and provenance are the
// reference's sensor interface (gpu/sensor.js:17-84 sensorParams, describe 87-92, analogGain 94-97, AutoExposure
// 127-155, rawNoiseDN 180-182, updateAWB 185-190); the code below is this module's own: one field table resolved
// against the physics file, the UVC exposure loop as a pure step on an exposure budget, and the two parameter
// objects the addon hands to vt_sensor and vt_isp (where the reference packs its 256-B WebGPU uniform,
// packSensorUniform 158-177, with the same derived numbers).

const PLANCK = 6.62607015e-34, LIGHT = 299792458;
const CHANNEL_NM = [600, 540, 460];                 // wavelengths standing for the R, G, B channels
export const CFA_CODES = Object.freeze({ RGGB: 0, GRBG: 1, GBRG: 2, BGGR: 3 });
export const ANALOG_GAIN_MAX = 8;                   // UVC gain 0..100 onto 1..8x (assumed, as the reference)

// [result key, provenance key, path in the physics file, default]; a leaf is a number (array, string) or a
// {value, status, source} box. The order is the result's field order.
const FIELDS = [
  ['pixelPitch', 'pixel_pitch', 'sensor.pixel_pitch', 3e-6],
  ['qePeak', 'qe_peak', 'sensor.qe_peak', [0.45, 0.55, 0.5]],
  ['fNumber', 'f_number', 'sensor.lens.f_number', 2.0],
  ['transmission', 'transmission', 'sensor.lens.transmission', 0.9],
  ['pupilExponent', 'pupil_obliquity_exponent', 'sensor.lens.relative_illumination.pupil_obliquity_exponent', 1.0],
  ['vignettingThetaDeg', 'vignetting_theta_deg', 'sensor.lens.relative_illumination.vignetting_theta_deg', [0, 90]],
  ['vignetting', 'vignetting', 'sensor.lens.relative_illumination.vignetting', [1, 1]],
  ['esfSigmaPx', 'esf_sigma_px', 'sensor.lens.esf_sigma_px', 0.7],
  ['readNoise', 'read_noise', 'sensor.read_noise', 3.0],
  ['fullWell', 'full_well', 'sensor.full_well', 10000],
  ['darkCurrent', 'dark_current', 'sensor.dark_current', 0],
  ['conversionGain', 'conversion_gain_base', 'sensor.conversion_gain_base', 0.33],
  ['adcBits', 'adc_bits', 'sensor.adc_bits', 12],
  ['blackLevel', 'black_level', 'sensor.black_level', 0],
  ['cfa', 'bayer_order', 'sensor.bayer_order', 'GRBG'],
  ['exposureUnit', 'exposure_unit', 'sensor.uvc.exposure_unit', 1e-4],
  ['exposureRange', 'exposure_range', 'sensor.uvc.exposure_range', [1, 5000]],
  ['gainRange', 'gain_range', 'sensor.uvc.gain_range', [0, 100]],
  ['fps', 'frame_rate', 'sensor.uvc.frame_rate', 20],
  ['aeTarget', 'target_mean', 'sensor.auto_exposure.target_mean', 0.45],
  ['aeDamping', 'damping', 'sensor.auto_exposure.damping', 0.3],
  ['aeMaxStep', 'max_step', 'sensor.auto_exposure.max_step', 2.0],
  ['wb', 'white_balance_gains', 'isp.white_balance_gains', [1, 1, 1]],
  ['ccm', 'ccm', 'isp.ccm', [[1, 0, 0], [0, 1, 0], [0, 0, 1]]],
  ['tone', 'tone_curve', 'isp.tone_curve', 'srgb'],
  ['sCurveC', 's_curve_c', 'isp.s_curve_c', null],
  ['blackOffset', 'black_offset_linear', 'isp.black_offset_linear', 0],
  ['demosaic', 'demosaic', 'isp.demosaic', 'bilinear'],
  ['noiseReduction', 'noise_reduction', 'isp.noise_reduction', 'none'],
  ['nrSigmaPx', 'noise_reduction_sigma_px', 'isp.noise_reduction_sigma_px', 0.4],
  ['glareFraction', 'veiling_glare.fraction', 'sensor.lens.veiling_glare.fraction', 0],
  ['glareSigmaFrac', 'veiling_glare.sigma_frac_diag', 'sensor.lens.veiling_glare.sigma_frac_diag', 0.25],
  ['gamma', 'gamma', 'isp.gamma', 2.2],
  ['sharpenAmount', 'sharpening.amount', 'isp.sharpening.amount', 0],
  ['sharpenRadius', 'sharpening.radius_px', 'isp.sharpening.radius_px', 1],
  ['chroma', 'chroma_subsampling', 'isp.chroma_subsampling', '4:2:0'],
  ['awb', 'white_balance_auto_default', 'sensor.uvc.white_balance_auto_default', true],
  ['analogGainRange', 'analog_gain_range', 'sensor.analog_gain_range', [1, ANALOG_GAIN_MAX]],
];
// constants of the model that no physics field sets
const ASSUMED = {
  nrRangeScale: [2.0, 'bilateral range sigma = 2 x the modelled noise'],
  glareLevel: [16, 'glare on a 1/16 pyramid level'],
  awbDamping: [0.3, 'grey-world AWB on the raw channel means, damped 0.3 per frame'],
};
const PROVENANCE_NOTES = {
  gainMapping: 'UVC gain 0..100 -> analog_gain_range, linear',
  awbMethod: 'grey-world AWB on the raw channel means, damped 0.3 per frame',
};

const isBox = (x) => x !== null && typeof x === 'object' && !Array.isArray(x) && 'value' in x;

/** The leaf at a dotted path, or undefined. */
function leaf(obj, path) {
  let x = obj;
  for (const k of path.split('.')) {
    if (x === null || x === undefined || typeof x !== 'object') return undefined;
    x = x[k];
  }
  return x;
}

/** Sensor + ISP parameters from physics.yaml's `sensor` and `isp` blocks ({value, status, source} leaves or plain
 * numbers), with each field's provenance; `over` replaces any field of the result (gpu/sensor.js:17-84). */
export function sensorParams(physics = {}, over = {}) {
  const p = {}, provenance = {};
  for (const [key, prov, path, dflt] of FIELDS) {
    const x = leaf(physics, path);
    const v = isBox(x) ? x.value : x;
    p[key] = v === undefined || v === null ? dflt : v;
    provenance[prov] = isBox(x) ? { status: x.status, source: x.source } : { status: 'default', source: 'js/sensor.mjs' };
  }
  p.responseMatrix = responseMatrix(physics, p, provenance);
  p.esfIncludesAperture = true;
  for (const [k, [v, why]] of Object.entries(ASSUMED)) {
    p[k] = v;
    provenance[k] = { status: 'assumed', source: why };
  }
  for (const [k, why] of Object.entries(PROVENANCE_NOTES)) provenance[k] = { status: 'assumed', source: why };
  p.awb = !!p.awb;
  p.noise = true;
  p.analogGainMax = (p.analogGainRange || [1, ANALOG_GAIN_MAX])[1];
  const e0 = leaf(physics, 'sensor.uvc.exposure_default');
  const e0v = isBox(e0) ? e0.value : e0;
  p.exposureUnits0 = e0v === undefined || e0v === null ? 10 : e0v;
  Object.assign(p, over);
  p.provenance = provenance;
  return p;
}

/** The response matrix M (e-/s per W m^-2) of the file, or, when it is missing, tagged a placeholder or too small to
 * be one, the diagonal of pixel area x QE x photons per joule at each channel's wavelength. */
function responseMatrix(physics, p, provenance) {
  const box = leaf(physics, 'sensor.response_matrix.M');
  const M = isBox(box) ? box.value : box;
  provenance.response_matrix = isBox(box) ? { status: box.status, source: box.source }
    : { status: 'default', source: 'js/sensor.mjs' };
  const placeholder = /placeholder/i.test(String((isBox(box) && box.source) || ''));
  const usable = Array.isArray(M) && !placeholder && M.flat().some((v) => Math.abs(v) >= 1e3);
  if (usable) return M;
  provenance.response_matrix = { status: 'derived', source: 'js/sensor.mjs: pixel_pitch^2 x qe_peak x lambda/(h c)' };
  const diag = CHANNEL_NM.map((nm, c) => (p.pixelPitch * p.pixelPitch * p.qePeak[c] * nm * 1e-9) / (PLANCK * LIGHT));
  return diag.map((g, r) => diag.map((_, c) => (r === c ? g : 0)));
}

/** Counts of the parameters by status, and the provenance table (gpu/sensor.js:87-92). */
export function describe(p) {
  const counts = {};
  for (const { status } of Object.values(p.provenance)) counts[status || 'unknown'] = (counts[status || 'unknown'] || 0) + 1;
  return { counts, provenance: p.provenance };
}

/** UVC gain units -> analogue gain: gainRange mapped linearly onto 1..analogGainMax. */
export function analogGain(p, uvcGain) {
  const [lo, hi] = p.gainRange;
  const t = (Math.min(hi, Math.max(lo, uvcGain)) - lo) / Math.max(1, hi - lo);
  return 1 + t * (p.analogGainMax - 1);
}

/** The exposure budget t x g split as a UVC camera does: whole exposure units first, up to the frame time, then
 * analogue gain for the rest; returns {units, gain, gainUnits, budget} with the budget clamped to what the pair can
 * reach. */
export function splitExposure(p, budget) {
  const unit = p.exposureUnit;
  const frameUnits = Math.min(p.exposureRange[1], Math.floor(1 / (p.fps * unit)));
  const units = Math.max(p.exposureRange[0], Math.min(frameUnits, Math.round(budget / unit)));
  const gain = Math.max(1, Math.min(p.analogGainMax, budget / (units * unit)));
  const [lo, hi] = p.gainRange;
  const gainUnits = Math.round(lo + ((gain - 1) / (p.analogGainMax - 1)) * (hi - lo));
  const reach = [p.exposureRange[0] * unit, frameUnits * unit * p.analogGainMax];
  return { units, gain, gainUnits, budget: Math.max(reach[0], Math.min(reach[1], budget)) };
}

/** One step of the exposure loop on the budget: the luma error, linearised through the tone curve (gamma 2.2,
 * or 1 for a linear tone), limited to 8x either way, damped (aeDamping as a power) and limited to aeMaxStep per
 * frame. */
export function exposureStep(p, budget, meanLuma) {
  const gamma = p.tone === 'linear' ? 1 : 2.2;
  const err = meanLuma > 1e-4 ? Math.pow(p.aeTarget / meanLuma, gamma) : 8;
  const step = Math.pow(Math.min(8, Math.max(0.125, err)), p.aeDamping);
  const lim = p.aeMaxStep || 2;
  return budget * Math.min(lim, Math.max(1 / lim, step));
}

/** UVC-style auto exposure (gpu/sensor.js:127-155) over exposureStep and splitExposure. update(meanLuma) takes the
 * last frame's mean output luma in 0..1 and returns the next {exposure, gain}. */
export class AutoExposure {
  constructor(p, { exposureUnits = 10, gainUnits = 0, enabled = true } = {}) {
    this.p = p;
    this.enabled = enabled;
    this.units = exposureUnits;
    this.gainUnits = gainUnits;
    this.budget = exposureUnits * p.exposureUnit * analogGain(p, gainUnits);
  }
  get exposure() { return this.units * this.p.exposureUnit; }
  get gain() { return analogGain(this.p, this.gainUnits); }
  update(meanLuma) {
    if (this.enabled) {
      const s = splitExposure(this.p, exposureStep(this.p, this.budget, meanLuma));
      this.units = s.units;
      this.gainUnits = s.gainUnits;
      this.budget = s.budget;
    }
    return { exposure: this.exposure, gain: this.gain };
  }
}

/** Modelled noise std in DN of one raw pixel at a mean of `e` electrons: shot and read noise in quadrature, times
 * the conversion gain and the analogue gain. */
export function rawNoiseDN(p, e, gain) {
  return Math.hypot(Math.sqrt(Math.max(0, e)), p.readNoise) * p.conversionGain * gain;
}

/** Grey-world white balance: the gains that would equalise the raw channel means to green, approached
 * geometrically, g_next = g^(1 - damping) target^damping per channel; unchanged while a channel is dark. */
export function updateAWB(prev, meanRGB, damping = 0.3) {
  const g = prev || [1, 1, 1];
  if (!meanRGB.every((m) => m > 1e-4)) return g;
  const target = meanRGB.map((m) => meanRGB[1] / m);
  return g.map((x, c) => Math.exp((1 - damping) * Math.log(x) + damping * Math.log(target[c])));
}

/** The raw stage's parameters for the addon (vt_sensor_params): the optics constant pi T / (4 N^2), the ESF sigma
 * left after the pixel aperture's 1/12, 2^bits - 1, the glare sigma on the 1/F level. */
export function nativeSensor(p, { width, height, frame = 0, seed = 0, exposure, gain }) {
  const esf = p.esfIncludesAperture ? Math.sqrt(Math.max(0, p.esfSigmaPx ** 2 - 1 / 12)) : p.esfSigmaPx;
  const F = p.glareLevel || 16;
  const cfa = CFA_CODES[p.cfa];
  return {
    cfa: cfa === undefined ? 1 : cfa, frame: frame >>> 0, seed: seed >>> 0,
    opticsK: (Math.PI * p.transmission) / (4 * p.fNumber * p.fNumber), esfSigma: esf, exposure, gain,
    m: p.responseMatrix.flat(), readNoise: p.readNoise, fullWell: p.fullWell, dark: p.darkCurrent,
    convGain: p.conversionGain, maxDn: 2 ** p.adcBits - 1, black: p.blackLevel, noise: p.noise ? 1 : 0,
    pupilExp: p.pupilExponent, glareFraction: p.glareFraction || 0,
    glareSigma: (p.glareSigmaFrac * Math.hypot(width, height)) / F, glareLevel: F,
    vignettingThetaDeg: Array.from(p.vignettingThetaDeg || []), vignetting: Array.from(p.vignetting || []),
  };
}

// the reference's ISP mode strings as the kernels' codes; the noise-reduction test is the reference's own regular
// expression (gpu/sensor.js:171), under which 'none' enables the filter too
const DEMOSAIC_CODE = (s) => (/hamilton/i.test(s) ? 2 : /edge/.test(s) ? 1 : 0);
const TONE_CODE = (tone, sCurve) => (tone === 'srgb' ? (sCurve ? 3 : 0) : tone === 'linear' ? 2 : 1);

/** The ISP's parameters for the addon (vt_isp_params). */
export function nativeIsp(p) {
  return {
    demosaic: DEMOSAIC_CODE(String(p.demosaic)),
    nr: /edge|bilateral|on/.test(String(p.noiseReduction)) ? 1 : 0,
    tone: TONE_CODE(p.tone, p.sCurveC), chroma420: String(p.chroma).startsWith('4:2:0') ? 1 : 0,
    wb: Array.from(p.wb), ccm: p.ccm.flat(), gamma: p.gamma, sCurveC: p.sCurveC || 1, blackOffset: p.blackOffset || 0,
    nrSigmaPx: p.nrSigmaPx || 0, nrRangeScale: p.nrRangeScale || 2, sharpenAmount: p.sharpenAmount,
    sharpenRadius: p.sharpenRadius,
  };
}
