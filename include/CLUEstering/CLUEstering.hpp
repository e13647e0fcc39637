// Umbrella header of the drop-in C++20 API (reference: CLUEstering/CLUEstering.hpp).
// Link with libclue_hip.so (the HIP engine behind the C ABI in clue_abi.h).
#pragma once
#include "CLUEstering/core/Clusterer.hpp"
#include "CLUEstering/core/ConvolutionalKernel.hpp"
#include "CLUEstering/core/Device.hpp"
#include "CLUEstering/core/DistanceMetrics.hpp"
#include "CLUEstering/data_structures/AssociationMap.hpp"
#include "CLUEstering/data_structures/DeviceAssociationMap.hpp"
#include "CLUEstering/data_structures/PointsDevice.hpp"
#include "CLUEstering/data_structures/PointsHost.hpp"
#include "CLUEstering/detail/Dim.hpp"
#include "CLUEstering/utils/cluster_centroid.hpp"
#include "CLUEstering/utils/get_clusters.hpp"
#include "CLUEstering/utils/read_csv.hpp"
#include "CLUEstering/utils/scores.hpp"
