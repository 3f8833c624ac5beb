// dropin_test.cpp -- the reference's bootstrapping tests, written against the
// reference's own function names, run on the MI355X engine through the drop-in
// header.  Build with -DTORUS32 or -DTORUS56.  Exit status 0 when every check
// passes.  Mirrors BOOTSTRAPPING.MP21_FR (bootstrappping_test.cpp:133-180),
// COMPARE.MP21_NTT's fused-vs-assembled agreement (compare_test.cpp:18-69),
// KSK.KSK (keyswitch_test.cpp:12-40) and BLIND_ROT.BLIND_ROT_WWL24_NTT
// (blindrot_test.cpp:279-333, TORUS32 defaults).
#include <cstdio>
#include <cstdlib>
#include <string>
#include "yatfhe_dropin.h"

static int failures = 0;
#define EXPECT_EQ(a, b, what)                                                                       \
    do {                                                                                            \
        if ((a) != (b)) {                                                                           \
            std::printf("FAIL %s: %lld != %lld\n", what, (long long)(a), (long long)(b));           \
            failures++;                                                                             \
        }                                                                                           \
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

static void mp21_fr(const std::string& which) {
    YatfheParameters param = params(which);
    param.torusBase = 8;
    initYatfhe(param);
    std::printf("%s: n:%d, k:%d, N:%d, b:%d, l:%d\n", which.c_str(), param.n, param.k, param.N, param.radixBits, param.l);
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
        blindRotateJP22Ntt(accum, bskMP, inputModN2, param);
        extractTlweFromTrlwe(tmp, accum, 0);
        switchKeyForTlwe(output, ksKey, tmp, param);
        const int decAft = symDecTlweToInt(output, tlweKey, p);
        std::printf("m: %d, slot: %d, decAft: %d\n", plain, slot, decAft);
        EXPECT_EQ(plain, decAft, "MP21_FR decrypted slot");
        // the fused entry point gives the same ciphertext, bit for bit
        Tlwe fused{param.n};
        functionalBootstrappingMPNtt(fused, input, bskMP, ksKey, v, param);
        for (int i = 0; i < param.n; i++) EXPECT_EQ(fused.a[i], output.a[i], "fused vs assembled mask");
        EXPECT_EQ(fused.b, output.b, "fused vs assembled body");
    }
    // the classic key (genBootstrappingKey) through functionalBootstrappingNtt
    BootstrappingKey bskClassic{param, param.lApprox};
    genBootstrappingKey(bskClassic, trgswKey, tlweKey, param);
    for (int plain = -p / 2; plain < p / 2; plain++) {
        const int slot = ((plain % p) + p) % p;
        Tlwe input{param.n};
        Tlwe output{param.n};
        symEncTlwe(input, modSwitchToTorusGeneral(slot, encMod, param.qLwe), tlweKey);
        functionalBootstrappingNtt(output, input, bskClassic, ksKey, v, param);
        EXPECT_EQ(plain, symDecTlweToInt(output, tlweKey, p), "classic key decrypted slot");
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
                param.lApprox, bsk.s2Level);
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

int main(int argc, char** argv) {
    setYatfheSeed(2024);
    try {
        mp21_fr("small");
#if defined(TORUS32)
        // the reference's TORUS32 defaults: n=680, N=2048, B=2^4, lApprox=4, KS 2^2 x 16
        if (argc < 2 || std::string(argv[1]) != "--small-only") {
            mp21_fr("default");
            wwl24_fr("default");
        }
#endif
    } catch (const std::exception& e) {
        std::printf("FAIL exception: %s\n", e.what());
        return 2;
    }
    std::printf("%s: %d failures\n", failures ? "FAILED" : "PASSED", failures);
    return failures ? 1 : 0;
}
