// GPU contact detector for triangle meshes (the Embree-based ClothContactDetector replaced by
// linear BVHs): Morton-sorted LBVH build, per-step refit, VF and EE discrete proximity queries,
// per-primitive minimum distances and per-vertex conservative bounds.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace gvbd {

// One BVH node: box lo/hi; lo.w (int bits) holds an internal node's left child or a leaf's
// primitive id, hi.w the node's skip link (next node in depth-first order after its subtree,
// -1 at the end). Internal nodes occupy [0, n-1), leaves [n-1, 2n-1). An edge leaf's box is
// exactly its two endpoints' coordinates, and bits 28-30 of its lo.w say, per axis, whether the
// edge's first vertex holds the hi coordinate: the leaf alone gives both endpoints, bit for bit.
constexpr int kLeafPrimBits = 28;
constexpr int kLeafPrimMask = (1 << kLeafPrimBits) - 1;
struct alignas(16) BvhNode {
    float4 lo;
    float4 hi;
};

// A VF contact: the face whose closest point to the vertex lies on primitive (type, id)
// (type 0 vertex, 1 edge, 2 face interior), d = distance.
struct alignas(16) VFContactRec {
    int32_t face;
    int32_t type;
    int32_t prim;
    float d;
};

// An EE contact: the other edge, the closest-point parameters on both edges and the distance.
struct alignas(16) EEContactRec {
    int32_t edge;
    float mu1;
    float mu2;
    float d;
};

struct DeviceBvh {
    int n = 0;                       // primitives
    BvhNode* nodes = nullptr;        // 2n - 1
    int* parent = nullptr;           // 2n - 1
    int* right = nullptr;            // n - 1: right child of each internal node (refit)
    int* jump = nullptr;             // 4 (n - 1): depth / ancestor buffers of the depth computation
    int* levelNodes = nullptr;       // internal nodes ordered by depth, deepest first
    std::vector<int> levelBegin;     // host: ranges of levelNodes per depth (deepest first)
    int* dLevelBegin = nullptr;      // the same ranges on the device
    std::vector<std::pair<int, int>> segments;  // refit launches: [first, last) level runs
    uint64_t* keys = nullptr;        // n (sorted Morton keys)
    uint64_t* keysAlt = nullptr;     // n
    uint64_t* depthKeys = nullptr;   // n - 1
    uint64_t* depthKeysAlt = nullptr;
    int* depth = nullptr;            // n - 1
    void* sortTemp = nullptr;
    size_t sortTempBytes = 0;
    int* mesh = nullptr;             // 2n - 1 (face trees of penetration contact): the mesh of every
                                     // face below the node, -1 where faces of several meshes meet
};

// Penetration contact of tet meshes (the reference tet solver's DCD and CCD,
// CollisionDetertionParameters.h:43-80, VBDPhysicsParameters.h:70).
struct PenetrationParams {
    int dcd, ccd;              // allowDCD, allowCCD
    int self;                  // handleSelfCollision
    int feasible;              // checkFeasibleRegion
    int traverse;              // checkTetTraverse
    int traverseNonSelf;       // tetrahedralTraverseForNonSelfIntersection
    int shiftToCenter;         // shiftQueryPointToCenter
    int stopAfterPassing;      // stopTraversingAfterPassingQueryPoint
    float offHeight;           // collisionOffHeight
    float feasibleEps;         // feasibleRegionEpsilon
    float rayEps;              // rayTriIntersectionEPSILON
    float maxSearchMult;       // maxSearchDistanceMultiplier
    float centerShift;         // centerShiftLevel
};

// The tets of a scene's tet meshes in scene ids (meshes in order, so a lower tet id means a
// lower (mesh, tet)), with the boundary surface's topology: the detector's faces are the meshes'
// boundary faces in the same order.
struct PenetrationSetup {
    std::vector<int32_t> tets;       // 4 per tet (scene vertex ids)
    std::vector<int32_t> tetMesh;    // per tet
    std::vector<int32_t> tetNbr;     // 4 per tet: scene tet across local face k (opposite vertex k), -1 outside
    std::vector<int32_t> faceOwner;  // 2 per detector face: scene tet and local face
    std::vector<int32_t> faceNbr;    // 3 per detector face: face across (v0 v1), (v1 v2), (v2 v0), -1 none
    std::vector<int32_t> vNbrBegin, vNbr;  // per scene vertex: its neighbours over surface edges
    std::vector<int32_t> surfVerts;  // boundary-surface vertices, ascending
    PenetrationParams params{};
};

// counts of the last detection pass
struct PenetrationCounts {
    int dcdHits, penetratedNoFace, ccdHits, walkOverflows, fullSearches;
};

struct DetectorResultCounts {
    int64_t vfPairs, vfContacts, eeContacts, vfOverflow, eeOverflow;
};

// Device pointers a solver reads after a detection (all sized by the detector's counts).
struct DetectorDeviceView {
    const int4* faces;
    const int4* edges;
    const int* vfStored;
    const VFContactRec* vfRec;
    int vfCap;
    const int* eeStored;
    const EEContactRec* eeRec;
    int eeCap;
    const int2* fvRange;
    const int* fvVerts;
    const float* bounds;
    const float4* xAtDetection;
    int* flag;
};

class ContactDetector {
  public:
    ContactDetector(int device, hipStream_t stream, float maxQueryDis, int vfCap, int eeCap);
    ~ContactDetector();
    ContactDetector(const ContactDetector&) = delete;
    ContactDetector& operator=(const ContactDetector&) = delete;

    // queryMask (optional, per vertex): 0 = the vertex takes no VF queries and gets an unbounded
    // conservative bound (interior vertices of a tet mesh whose faces are its boundary surface)
    int addMesh(int nv, const float* pos, int nf, const int* faces, const uint8_t* queryMask = nullptr);
    // Queries read (and refits follow) positions owned by the caller, e.g. a scene's x array in
    // the same vertex order; call before finalize.
    void useExternalPositions(float4* pos);
    // the external positions moved (a scene copied out and back at another address)
    void rebindPositions(float4* pos);
    void finalize();
    // Face back-links of the stored VF contacts (ClothContactDetector::faceContactInfos): per
    // face the lowest `cap` vertex ids whose stored contacts name that face.
    void buildFaceLinks(int cap, const int* gate = nullptr);
    // positionsAtPrevCollisionDetection = x; clears the re-detection flag
    void snapshotPositions(const int* gate = nullptr);
    // One detection pass of the cloth contact model whose run is decided on the device: a gate
    // set when `first` (a substep's first iteration) or when the re-detection flag is set (a
    // vertex was truncated in the previous iteration); with the gate clear every kernel of the
    // pass returns at once and the detection state stays as it was. The same results as calling
    // updateBvh, query, bounds, buildFaceLinks and snapshotPositions when the host sees the flag
    // set, without a host read, so the pass can be captured in a graph. `rebuild` (first only)
    // rebuilds the trees instead of refitting them (a host read: not capturable).
    void detectGated(bool first, bool rebuild, float thickness, float relaxation);
    // passes run by detectGated since finalize (a host read)
    int64_t gatedPasses();
    // Penetration contact of tet meshes instead of the VF / EE proximity queries (call before
    // finalize). Per pass: penetrationDcd at the start of a substep, then penetrationCcd after the
    // initial step with the substep's start positions, then buildFaceLinks. Hits are written as
    // one VF record per vertex (face = the face the vertex is pushed out of), which the face links
    // and the solve read.
    void setPenetration(PenetrationSetup&& setup);
    bool penetration() const { return pen_; }
    // tree refit (full rebuild when `rebuild`) at the current positions, then DCD
    void penetrationDcd(bool rebuild);
    void penetrationCcd(const float4* prev, bool rebuild);
    PenetrationCounts penetrationCounts();
    // per vertex: hit face (-1 none) and flags (1 a tet contained it, 2 the hit came from CCD)
    void penetrationResults(int32_t* face, uint8_t* flags);
    DetectorDeviceView deviceView() const;
    bool finalized() const { return ready_; }

    void setPositions(int mesh, const float* pos);
    void getPositions(int mesh, float* pos) const;
    // x = rest + step * drift[mesh] + sigma * N(0, 1) (counter-based, seeded per (seed, step, vertex))
    void perturb(uint64_t seed, int step, float sigma, const float* drift);
    void updateBvh(bool rebuild);
    void query(bool record, const int* gate = nullptr);
    void bounds(float thickness, float relaxation, const int* gate = nullptr);

    DetectorResultCounts counts();
    // record-mode outputs (valid after query(true)); sizes from counts()
    void vfPairs(int32_t* vf, float* d);
    void vfContactsAll(int32_t* rec, float* d);
    void eeContactsAll(int32_t* ee, float* mu, float* d);
    // capped outputs (every query): per vertex vfCap, per edge eeCap records and the stored count
    void vfContactsCapped(int32_t* count, VFContactRec* rec);
    void eeContactsCapped(int32_t* count, EEContactRec* rec);
    void minDistances(float* vf, float* face, float* edge);
    void boundsOut(float* b);

    int numVertices() const { return nv_; }
    int numFaces() const { return nf_; }
    int numEdges() const { return ne_; }
    int numMeshes() const { return (int)meshV_.size() - 1; }
    void meshOffsets(int mesh, int* v, int* f, int* e) const;
    void topology(int32_t* edges, int32_t* faceEdges) const;
    size_t deviceBytes() const { return bytes_; }
    hipStream_t stream() const { return stream_; }

  private:
    // kind 0 faces, 1 edges, 3 precomputed primitive boxes (`boxes`: lo, hi per primitive)
    void buildBvh(DeviceBvh& b, int kind, const float4* boxes = nullptr);
    void refitBvh(DeviceBvh& b, int kind, hipStream_t stream, const float4* boxes = nullptr,
                  const int* gate = nullptr);
    void allocBvh(DeviceBvh& b, int n);
    void setupPenetrationDevice();
    // the face-side and edge-side work of a pass are independent: side_ runs the edge side
    // between fork() (side_ waits for stream_) and join() (stream_ waits for side_)
    void fork();
    void join();
    void freeAll();
    template <class T> T* dalloc(size_t n);

    int device_;
    hipStream_t stream_;
    hipStream_t side_ = nullptr;  // default priority, created at finalize
    hipEvent_t forkEv_ = nullptr, joinEv_ = nullptr;
    float r_;
    int vfCap_, eeCap_;
    bool ready_ = false;
    bool treeCurrent_ = false;  // the trees were updated after the last position upload / perturb
    size_t bytes_ = 0;
    std::vector<void*> allocs_;
    // host topology (global ids)
    int nv_ = 0, nf_ = 0, ne_ = 0;
    std::vector<int> meshV_, meshF_, meshE_;  // offsets per mesh (size M + 1 after finalize)
    std::vector<float> hostPos_;
    std::vector<int> hFaces_, hFaceEdges_, hEdges_, hVMesh_, hEMesh_;
    std::vector<uint8_t> hQuery_;  // per vertex: takes VF queries
    bool masked_ = false;          // some mesh came with a query mask
    uint8_t* vQuery_ = nullptr;    // device copy of hQuery_ when masked_
    // device
    float4 *pos_ = nullptr, *rest_ = nullptr;
    int *vMesh_ = nullptr, *eMesh_ = nullptr;
    int4 *faces_ = nullptr, *faceEdges_ = nullptr, *edges_ = nullptr;
    int *vvBegin_ = nullptr, *vvNei_ = nullptr, *vfBegin_ = nullptr, *vfFace_ = nullptr, *veBegin_ = nullptr,
        *veEdge_ = nullptr;
    float* drift_ = nullptr;
    DeviceBvh fb_, eb_;
    float* sceneBox_ = nullptr;  // 6 floats (as ordered ints during reduction)
    // outputs
    float *vfMin_ = nullptr, *eeMin_ = nullptr, *bounds_ = nullptr;
    int* faceMin_ = nullptr;  // float bits
    int *vfPairCount_ = nullptr, *vfStored_ = nullptr, *eeCount_ = nullptr, *eeStored_ = nullptr;
    int *vfOverflow_ = nullptr, *eeOverflow_ = nullptr;
    VFContactRec* vfRec_ = nullptr;
    EEContactRec* eeRec_ = nullptr;
    int64_t *vfOffset_ = nullptr, *eeOffset_ = nullptr;
    void* scanTemp_ = nullptr;
    size_t scanTempBytes_ = 0;
    // record mode buffers (grown on demand)
    int64_t nRecPairs_ = 0, nRecEE_ = 0, capPairs_ = 0, capEE_ = 0;
    int2* recVF_ = nullptr;
    float* recVFd_ = nullptr;
    int4* recKeys_ = nullptr;  // (v, type, prim, face), type -1 = unused slot
    float* recKeysD_ = nullptr;
    int2* recEE_ = nullptr;
    float2* recEEmu_ = nullptr;
    float* recEEd_ = nullptr;
    bool recorded_ = false;
    int* error_ = nullptr;      // bit 0: traversal stack overflow, bit 1: tree too deep
    int* depthHist_ = nullptr;  // per depth: first index of its run in the refit order (build)
    unsigned int* boxPartial_ = nullptr;  // per-block centroid bounds (build)
    int boxPartialBlocks_ = 0;            // blocks boxPartial_ holds
    float4* external_ = nullptr;
    // face back-links and the solver-side state
    int* fvVerts_ = nullptr;    // nv * vfCap: vertex ids grouped by face (FV lists)
    int* fvCounts_ = nullptr;   // 2 nf + 1: per-face counts (nf + 1), then scatter cursors (nf)
    int* fvOffsets_ = nullptr;  // nf + 1: exclusive scan of the counts
    int2* fvRange_ = nullptr;
    void* fvScanTemp_ = nullptr;
    uint64_t* vKeys_ = nullptr;     // vertex Morton keys (VF query order)
    uint64_t* vKeysAlt_ = nullptr;  // sorted
    void* vSortTemp_ = nullptr;
    size_t vSortTempBytes_ = 0;
    size_t fvScanTempBytes_ = 0;
    float4* xAtCD_ = nullptr;
    int* flag_ = nullptr;
    int* gate_ = nullptr;        // detectGated: whether the current pass runs
    int* gatePasses_ = nullptr;  // detectGated: passes run
    // penetration contact
    bool pen_ = false;
    PenetrationSetup penHost_;
    int nt_ = 0, nSurf_ = 0;
    int4 *penTets_ = nullptr, *penTetNbr_ = nullptr, *penFaceNbr_ = nullptr;
    int *penTetMesh_ = nullptr, *penVNbrBegin_ = nullptr, *penVNbr_ = nullptr, *penSurf_ = nullptr;
    int2* penFaceOwner_ = nullptr;
    float4 *tetBoxes_ = nullptr, *sweptBoxes_ = nullptr;
    uint8_t* penFlags_ = nullptr;
    int* penCounts_ = nullptr;  // kPenCounts: DCD hits, penetrated without a face, CCD hits, walk
                                // overflows, contained vertices listed by the DCD's first pass
    int* penInTet_ = nullptr;   // per vertex: its containing tet (vertices listed by the DCD)
    int* penList_ = nullptr;    // per surface vertex slot: the contained vertices of the DCD
    DeviceBvh tb_, sb_;         // tets, faces over their path in a substep
    bool sweptBuilt_ = false;
};

}  // namespace gvbd
