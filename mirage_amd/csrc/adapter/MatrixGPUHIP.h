// MatrixGPUHIP.h - the GPU engine entry points that MirAge's Analyser calls,
// as MatrixGPUHIP.cpp defines them.  The signatures are the reference's own
// (MirAge/Analyser.h:228-232, MirAge/Hardware.h:16); Analyser.h declares the
// first three only when __CUDA__ is defined, so this header declares them
// for a build or a test that leaves it undefined.  Include after Analyser.h.
#pragma once

vector<MatrixResult> CUDAGetHighestScore(
    vector<vector<char> *> seqQuery, Sequences *sequences, bool debug, int cudaDeviceID,
    size_t ConsiderThreshold, vector<SEQ_ID_t> *queryReferenceCombinations,
    vector<vector<SEQ_POS_t>> *queryReferenceCombinationsRefPositions,
    vector<vector<SEQ_POS_t>> *queryReferenceCombinationsQueryPositions,
    vector<vector<Enum::Position>> *queryReferenceCombinationsCombiPositions,
    vector<vector<Enum::Position>> *queryReferenceCombinationsCombiPositions_LastPositionInRange,
    int threadSize, Display *display, MultiThreading *multiThreading,
    SEQ_AMOUNT_t referencesToAnalyse_Offset, SEQ_AMOUNT_t referencesToAnalyse_SizePerAnalysis,
    SEQ_AMOUNT_t queriesToAnalyse_Offset, SEQ_AMOUNT_t queriesToAnalyse_SizePerAnalysis,
    vector<vector<Enum::HeadTailSequence>> *matchedReferencesPerQuery_HeadTailSequences);
void InitialiseCuda(int numDevices);
size_t Get_LOCAL_DATA_rowhit_SIZE();
size_t CUDAGetMaxBlockSize(int cudaDeviceID);

// Builds every initialised device's index from the reference sequences
// (Sequences::referenceSequences) with the run's considerthreshold.  No
// counterpart in the reference: its engine re-sends reference bytes with
// every call, this one keeps them resident in HBM.
void MirageHipLoadReferences(const vector<vector<char>> &refs, int32_t considerThreshold);
