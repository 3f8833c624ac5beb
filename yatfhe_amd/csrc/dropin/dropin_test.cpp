// dropin_test.cpp -- the reference's bootstrapping tests, written against the
// reference's own types, function names and call forms, run on the MI355X
// engine through the drop-in header.  Build with -DTORUS32 or -DTORUS56.  Exit
// status 0 when every check passes.  Mirrors
//   BOOTSTRAPPING.MP21_FR (bootstrappping_test.cpp:133-180), with the
//     reference's blindRotateMP21Ntt(accum, bsKey.bskDft, ...) call and the
//     fused functionalBootstrappingMPNtt agreeing bit for bit,
//   COMPARE.MP21_NTT (compare_test.cpp:18-69) verbatim in its calls: the
//     reference's default parameters, BootstrappingKeyMP bsKey{param} (the
//     full param.l levels), extraction at param.driftPhase != 0,
//   BLIND_ROT.BLIND_ROT_NTT's classic form blindRotateNtt(in2, bsk.bskDft, ...)
//     (blindrot_test.cpp:191) and functionalBootstrappingNtt,
//   KSK.KSK (keyswitch_test.cpp:12-40),
//   BLIND_ROT.BLIND_ROT_WWL24_NTT (blindrot_test.cpp:279-333),
// and the serving benchmark's concurrent workers (blindrotate_serving.cpp:
// 197-228): bootstraps from several threads at once equal the serial ones.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <thread>
#include "yatfhe_dropin.h"

static int failures = 0;
#define EXPECT_EQ(a, b, what)                                                                       \
    do {                                                                                            \
        if ((a) != (b)) {                                                                           \
            std::printf("FAIL %s: %lld != %lld\n", what, (long long)(a), (long long)(b));           \
            failures++;                                                                             \
        }                                                                                           \
    } while (0)
#define EXPECT_TRUE(c, what)                         \
    do {                                             \
        if (!(c)) {                                  \
            std::printf("FAIL %s\n", what);          \
            failures++;                              \
        }                                            \
    } while (0)

static YatfheParameters params(const std::string& which) {
    YatfheParameters p{};
    if (which == "small") {  // n=630, N=1024 (TORUS32) or n=742, N=2048 (TORUS56)
#if defined(TORUS32)
        p.n = 630; p.N = 1024; p.setRadixBits(7); p.l = 3; p.lApprox = 3;
#else
        p.n = 742; p.N = 2048; p.setRadixBits(10); p.l = 3; p.lApprox = 3;
#endif
        p.ksRadixBits = 2; p.ksWidthBits = 32; p.ksLevel = 8;
    }
    return p;  // "default": the reference's own defaults (yatfhe_parameters.h:30-77)
}

static bool same(const Tlwe& a, const Tlwe& b) {
    return a.a == b.a && a.b == b.b;
}

static void mp21_fr(const std::string& which) {
    YatfheParameters param = params(which);
    param.torusBase = 8;
    const YatfheParameters before = param;
    initYatfhe(param);
    EXPECT_EQ(param.l, before.l, "initYatfhe leaves param.l alone");
    std::printf("%s: n:%d, k:%d, N:%d, b:%d, l:%d, lApprox:%d\n", which.c_str(), param.n, param.k, param.N, param.radixBits, param.l,
                param.lApprox);
    const int p = param.torusBase;
    const int encMod = 2 * p;
    TlweKey tlweKey{param};
    TrgswKey trgswKey{param};
    TrlweKey& trlweKey = trgswKey.trlweKey;
    BootstrappingKeyMP bskMP{param, param.lApprox};
    TlweKeySwitchingKey ksKey{param};
    genTlweKey(tlweKey);
    genTrlweKey(trlweKey);
    genBootstrappingKeyMP(bskMP, trgswKey, tlweKey, param);
    TlweKey tlweKsKey = tlweKey;
    genTlweKeySwitchingKey(ksKey, trlweKey, tlweKsKey, param);
    TorusPolynomial v{param.N};
    generateTestPolynomialFR(v, p, 2 * param.N);
    std::vector<Tlwe> inputs, serial;
    for (int plain = -p / 2; plain < p / 2; plain++) {
        const int slot = ((plain % p) + p) % p;
        Torus mu = modSwitchToTorusGeneral(slot, encMod, param.qLwe);
        Tlwe input{param.n};
        Tlwe output{param.n};
        symEncTlwe(input, mu, tlweKey);
        ScaledTlwe inputModN2{2 * param.N, param.n};
        Trlwe accum{param};
        Tlwe tmp{ksKey.nCurrKey};
        rescaleTlweToNewMod(inputModN2, input);
        genNoiselessTrlweSample(accum, v, inputModN2);
        blindRotateMP21Ntt(accum, bskMP.bskDft, inputModN2, param);
        extractTlweFromTrlwe(tmp, accum, param.driftPhase);
        switchKeyForTlwe(output, ksKey, tmp, param);
        const int decAft = symDecTlweToInt(output, tlweKey, p);
        std::printf("m: %d, slot: %d, decAft: %d\n", plain, slot, decAft);
        EXPECT_EQ(plain, decAft, "MP21_FR decrypted slot");
        // blindRotateJP22Ntt on the key object computes the same accumulator
        Trlwe accum2{param};
        genNoiselessTrlweSample(accum2, v, inputModN2);
        blindRotateJP22Ntt(accum2, bskMP, inputModN2, param);
        EXPECT_TRUE(accum2.b.coeffs == accum.b.coeffs && accum2.a[0].coeffs == accum.a[0].coeffs, "JP22 == MP21 accumulator");
        // the fused entry point gives the same ciphertext, bit for bit
        Tlwe fused{param.n};
        functionalBootstrappingMPNtt(fused, input, bskMP, ksKey, v, param);
        EXPECT_TRUE(same(fused, output), "fused vs assembled bootstrap");
        inputs.push_back(input);
        serial.push_back(fused);
    }
    // the reference's serving workers: the same bootstraps from four threads at once
    {
        std::vector<Tlwe> conc(inputs.size(), Tlwe{param.n});
        std::vector<std::thread> ts;
        for (int t = 0; t < 4; t++)
            ts.emplace_back([&, t] {
                for (size_t i = t; i < inputs.size(); i += 4) functionalBootstrappingMPNtt(conc[i], inputs[i], bskMP, ksKey, v, param);
            });
        for (auto& th : ts) th.join();
        for (size_t i = 0; i < inputs.size(); i++) EXPECT_TRUE(same(conc[i], serial[i]), "concurrent == serial bootstrap");
    }
    // the classic key (genBootstrappingKey) through blindRotateNtt(accum, bsk.bskDft, ...)
    // and functionalBootstrappingNtt: computed for l == lApprox only
    YatfheParameters pc = param;
    pc.l = pc.lApprox;
    BootstrappingKey bskClassic{pc, pc.lApprox};
    genBootstrappingKey(bskClassic, trgswKey, tlweKey, pc);
    for (int plain = -p / 2; plain < p / 2; plain++) {
        const int slot = ((plain % p) + p) % p;
        Tlwe input{param.n};
        Tlwe output{param.n};
        symEncTlwe(input, modSwitchToTorusGeneral(slot, encMod, param.qLwe), tlweKey);
        functionalBootstrappingNtt(output, input, bskClassic, ksKey, v, pc);
        EXPECT_EQ(plain, symDecTlweToInt(output, tlweKey, p), "classic key decrypted slot");
        ScaledTlwe inputModN2{2 * param.N, param.n};
        Trlwe accum{param};
        Tlwe tmp{ksKey.nCurrKey}, out2{param.n};
        rescaleTlweToNewMod(inputModN2, input);
        genNoiselessTrlweSample(accum, v, inputModN2);
        blindRotateNtt(accum, bskClassic.bskDft, inputModN2, pc);
        extractTlweFromTrlwe(tmp, accum, pc.driftPhase);
        switchKeyForTlwe(out2, ksKey, tmp, pc);
        EXPECT_TRUE(same(out2, output), "classic assembled vs fused");
    }
    if (param.l != param.lApprox) {
        bool threw = false;
        try {
            Tlwe input{param.n}, output{param.n};
            symEncTlwe(input, 0, tlweKey);
            functionalBootstrappingNtt(output, input, bskClassic, ksKey, v, param);
        } catch (const std::invalid_argument&) {
            threw = true;
        }
        EXPECT_TRUE(threw, "classic path with l != lApprox throws std::invalid_argument");
    }
    // KSK.KSK: a fresh encryption under the flattened GLWE key switches to the LWE key
    TlweKey inKey(param.k * param.N, param.rlweNoiseB);
    for (int j = 0; j < param.N; j++) inKey.s[j] = trlweKey.s[0].coeffs[j];
    for (int m = -p / 2; m < p / 2; m++) {
        Tlwe in{param.k * param.N};
        Tlwe out{param.n};
        symEncTlwe(in, modSwitchToTorusGeneral(m, p, param.qLwe), inKey);
        switchKeyForTlwe(out, ksKey, in, param);
        EXPECT_EQ(m, symDecTlweToInt(out, tlweKey, p), "KSK.KSK message");
    }
}

// COMPARE.MP21_NTT (compare_test.cpp:18-69): the reference's defaults, a full
// param.l-level key, extraction at param.driftPhase = N / torusBase / 2.
static void compare_mp21_ntt() {
    YatfheParameters param{};
    constexpr int kPlainMod = 1 << 4;
    param.torusBase = kPlainMod;
    param.driftPhase = param.N / param.torusBase / 2;
    initYatfhe(param);
    std::printf("COMPARE.MP21_NTT n:%d, k:%d, N:%d, b:%d, l:%d, driftPhase:%d\n", param.n, param.k, param.N, param.radixBits, param.l,
                param.driftPhase);
    TlweKey tlweKey{param};
    TrgswKey trgswKey{param};
    TrlweKey& trlweKey = trgswKey.trlweKey;
    BootstrappingKeyMP bsKey{param};
    TlweKeySwitchingKey ksKey{param};
    genTlweKey(tlweKey);
    genTrlweKey(trlweKey);
    genBootstrappingKeyMP(bsKey, trgswKey, tlweKey, param);
    TlweKey tlweKsKey = tlweKey;
    genTlweKeySwitchingKey(ksKey, trlweKey, tlweKsKey, param);
    int threshold = kPlainMod / 4;
    TorusPolynomial v{param.N};
    generateTestPolynomialLt1(v, threshold);
    for (int plain = 0; plain < kPlainMod / 2; plain += 1) {
        Torus mu = modSwitchToTorusGeneral(plain, kPlainMod, param.qLwe);
        Tlwe input{param.n};
        Tlwe output{param.n};
        symEncTlwe(input, mu, tlweKey);
        ScaledTlwe inputModN2{param.N * 2, param.n};
        Trlwe accum{param.k, param.N};
        Tlwe tmp{ksKey.nCurrKey};
        rescaleTlweToNewMod(inputModN2, input);
        genNoiselessTrlweSample(accum, v, inputModN2);
        blindRotateMP21Ntt(accum, bsKey.bskDft, inputModN2, param);
        extractTlweFromTrlwe(tmp, accum, param.driftPhase);
        switchKeyForTlwe(output, ksKey, tmp, param);
        auto decAft = symDecTlweToInt(output, tlweKey, kPlainMod);
        const int expected = plain < threshold ? 1 : 0;
        std::printf("plain: %d, decAft: %d\n", plain, static_cast<int>(decAft));
        EXPECT_EQ(expected, decAft, "COMPARE.MP21_NTT");
        // the fused call extracts at param.driftPhase too
        Tlwe fused{param.n};
        functionalBootstrappingMPNtt(fused, input, bsKey, ksKey, v, param);
        EXPECT_TRUE(same(fused, output), "COMPARE.MP21_NTT fused vs assembled");
    }
}

// BLIND_ROT.BLIND_ROT_WWL24_NTT (blindrot_test.cpp:279-333) carried through the
// bootstrap: blindRotateWWL24Ntt with a succinct key decrypts every slot.
static void wwl24_fr(const std::string& which) {
    YatfheParameters param = params(which);
    param.torusBase = 8;
    initYatfhe(param);
    const int p = param.torusBase;
    TlweKey tlweKey{param};
    TrgswKey trgswKey{param};
    TrlweKey& trlweKey = trgswKey.trlweKey;
    BootstrappingKeyWWL24 bsk{param, param.lApprox};
    TlweKeySwitchingKey ksKey{param};
    genTlweKey(tlweKey);
    genTrlweKey(trlweKey);
    genBootstrappingKeyWWL24(bsk, trgswKey, tlweKey, param);
    genTlweKeySwitchingKey(ksKey, trlweKey, tlweKey, param);
    std::printf("%s WWL+24: n:%d, N:%d, b:%d, lApprox:%d, s2 levels:%d\n", which.c_str(), param.n, param.N, param.radixBits,
                param.lApprox, bsk.s2Dft.l);
    TorusPolynomial v{param.N};
    generateTestPolynomialFR(v, p, 2 * param.N);
    for (int plain = -p / 2; plain < p / 2; plain++) {
        const int slot = ((plain % p) + p) % p;
        Tlwe input{param.n};
        Tlwe output{param.n};
        symEncTlwe(input, modSwitchToTorusGeneral(slot, 2 * p, param.qLwe), tlweKey);
        ScaledTlwe inputModN2{2 * param.N, param.n};
        Trlwe accum{param};
        Tlwe tmp{ksKey.nCurrKey};
        rescaleTlweToNewMod(inputModN2, input);
        genNoiselessTrlweSample(accum, v, inputModN2);
        blindRotateWWL24Ntt(accum, bsk, inputModN2, param);
        extractTlweFromTrlwe(tmp, accum, 0);
        switchKeyForTlwe(output, ksKey, tmp, param);
        const int decAft = symDecTlweToInt(output, tlweKey, p);
        std::printf("m: %d, slot: %d, decAft: %d\n", plain, slot, decAft);
        EXPECT_EQ(plain, decAft, "WWL24 decrypted slot");
    }
}

// blindRotateLazyPipeNtt and blindRotateLazyPipeAltNtt (the serving
// benchmark's OURS keys) with the reference's TORUS32 lazy-pipe gadget (2^4,
// l = 8, lApprox = 4) at N = 1024: every slot decrypts after extraction and
// keyswitch, with the reference call forms (blind_rotate.h:38-46).
static void lazy_fr() {
#if defined(TORUS32)
    YatfheParameters param{};
    param.n = 630; param.N = 1024; param.setRadixBits(4); param.l = 8; param.lApprox = 4;
    param.torusBase = 4;
    initYatfhe(param);
    const int p = param.torusBase;
    TlweKey tlweKey{param};
    TrgswKey trgswKey{param};
    TrlweKey& trlweKey = trgswKey.trlweKey;
    TlweKeySwitchingKey ksKey{param};
    genTlweKey(tlweKey);
    genTrlweKey(trlweKey);
    genTlweKeySwitchingKey(ksKey, trlweKey, tlweKey, param);
    TorusPolynomial v{param.N};
    generateTestPolynomialFR(v, p, 2 * param.N);
    BootstrappingKeyMPLazyPipe pipe{param, param.lApprox, true, false};
    genBootstrappingKeyMPLazyPipe(pipe, trgswKey, tlweKey, v, param);
    BootstrappingKeyMPLazyPipeAlt alt{param, param.lApprox, true};
    genBootstrappingKeyMPLazyPipeAlt(alt, trgswKey, tlweKey, param);
    std::vector<NttPolynomial> gdVntt;
    prepareGdV(gdVntt, v, param);
    std::printf("lazy pipe: n:%d, N:%d, b:%d, l:%d, lApprox:%d\n", param.n, param.N, param.radixBits, param.l, param.lApprox);
    for (int plain = -p / 2; plain < p / 2; plain++) {
        const int slot = ((plain % p) + p) % p;
        Tlwe input{param.n};
        symEncTlwe(input, modSwitchToTorusGeneral(slot, 2 * p, param.qLwe), tlweKey);
        ScaledTlwe scaled{2 * param.N, param.n};
        rescaleTlweToNewMod(scaled, input);
        for (int form = 0; form < 2; form++) {
            Trlwe accum{param};
            if (form == 0) blindRotateLazyPipeNtt(accum, pipe, scaled, v, param);
            else blindRotateLazyPipeAltNtt(accum, alt, scaled, v, gdVntt, param);
            Tlwe tmp{ksKey.nCurrKey};
            Tlwe output{param.n};
            extractTlweFromTrlwe(tmp, accum, 0);
            switchKeyForTlwe(output, ksKey, tmp, param);
            const int dec = symDecTlweToInt(output, tlweKey, p);
            std::printf("%s m: %d, slot: %d, dec: %d\n", form ? "alt" : "pipe", plain, slot, dec);
            EXPECT_EQ(plain, dec, form ? "LazyPipeAlt decrypted slot" : "LazyPipe decrypted slot");
        }
    }
#endif
}

int main(int argc, char** argv) {
    setYatfheSeed(2024);
    try {
        mp21_fr("small");
        if (argc < 2 || std::string(argv[1]) != "--small-only") {
            compare_mp21_ntt();
#if defined(TORUS32)
            // the reference's TORUS32 defaults: n=680, N=2048, B=2^4, l=8, lApprox=4, KS 2^2 x 16
            mp21_fr("default");
            wwl24_fr("default");
            lazy_fr();
#endif
        }
    } catch (const std::exception& e) {
        std::printf("FAIL exception: %s\n", e.what());
        return 2;
    }
    std::printf("%s: %d failures\n", failures ? "FAILED" : "PASSED", failures);
    return failures ? 1 : 0;
}
