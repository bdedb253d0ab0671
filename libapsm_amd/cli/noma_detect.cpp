// NOMA_detect: trains a detector on the pilots of a synchronized data file and detects the
// data symbols (cpp/cli/detect/noma_detect.cpp:65-153, flags in cpp/TOOL_overview.md:76-96).
//
// -p (all users) runs one multi-user APSM detector: the users share the received pilots, so
// they share the dictionary and every kernel evaluation, where the reference starts one
// detector per user in its own host thread.
#include <cstdlib>
#include <iostream>
#include <memory>
#include <string>

#include "cli_common.hpp"
#include "noma/noma_detector_factory.cuh"

namespace {

int run( int argc, char** argv )
{
    cli::ArgParser args( "NOMA_detect", "0.0.3",
                         "This tool runs the training and detection algorithms. The algorithm configuration is passed in JSON format "
                         "either as string or in a file." );
    args.add( "-s", "--synchronized-data", "synchronized data file containing received training and data signals", "", false, true );
    args.add( "-r", "--reference-data", "reference data file" );
    args.add( "-rs", "--random-seed", "random seed for generation of data (this overrides the --reference-data argument)" );
    args.add( "-d", "--detected-data", "detected (estimated) data file" );
    args.add( "-cf", "--config-file", "filename for JSON file containing the algorithm configuration" );
    args.add( "-cs", "--config-string", "JSON string containing the algorithm configuration", "{ \"algorithm\": { \"otype\": \"APSM\" } }" );
    args.add( "-o", "--output-mode",
              "print mode for error metric (BER = bit errors, SER = symbol errors, EVM = Error Vector Magnitude, scripting = disable all "
              "output except \"ber,ser,evm\")",
              "SER" );
    args.add( "-th", "--train-history-file", "filename for training history CSV-file (test set will be used as validation set)" );
    args.add( "-p", "--parallel-processing", "process all users in parallel", "false", true );
    try
    {
        if ( !args.parse( argc, argv ) )
            return EXIT_SUCCESS;
    }
    catch ( const std::runtime_error& e )
    {
        std::cerr << e.what() << std::endl << args.usage();
        return EXIT_FAILURE;
    }

    // configuration (argparse_helper.cpp:234-257)
    nlohmann::json config;
    if ( args.present( "--config-file" ) )
    {
        std::ifstream f( args.get( "--config-file" ) );
        if ( !f )
            throw std::runtime_error( "cannot open config file " + args.get( "--config-file" ) );
        config = nlohmann::json::parse( f );
    }
    else
        config = nlohmann::json::parse( args.get( "--config-string" ) );

    // output mode (argparse_helper.cpp:209-232)
    std::string om = args.get( "--output-mode" );
    std::transform( om.begin(), om.end(), om.begin(), ::toupper );
    cli::OutputMode mode;
    if ( om == "BER" )
        mode = cli::BER;
    else if ( om == "SER" )
        mode = cli::SER;
    else if ( om == "EVM" )
        mode = cli::EVM;
    else if ( om == "SCRIPTING" )
    {
        std::cout.setstate( std::ios_base::failbit );
        mode = cli::Scripting;
    }
    else
        throw std::invalid_argument( "Invalid print mode: " + args.get( "--output-mode" ) );

    // received signals (argparse_helper.cpp:160-178)
    const std::string synced = args.get( "--synchronized-data" );
    ComplexSampleMatrix rxTrain = cli::loadComplex( synced, "rxSigTraining" );
    ComplexSampleMatrix rxData = cli::loadComplex( synced, "rxSigData" );

    // reference symbols (argparse_helper.cpp:77-105)
    const cli::Modulation mod = cli::Modulation::fromString( config.value( "modulation", "QAM16" ) );
    ComplexSampleMatrix txTrain, txData;
    if ( args.present( "--random-seed" ) )
        cli::generateReferenceData( txTrain, txData, (uint32_t)std::stoul( args.get( "--random-seed" ) ), mod );
    else if ( args.present( "--reference-data" ) )
    {
        txTrain = cli::loadComplex( args.get( "--reference-data" ), "mod_training", mod.getScale() );
        txData = cli::loadComplex( args.get( "--reference-data" ), "mod_data", mod.getScale() );
    }
    else
        throw std::invalid_argument( "Neither --random-seed/-rs nor --reference-data/-r was specified" );

    const uint32_t user = config.value( "user", 0 );
    nlohmann::json antennas = nlohmann::json { { "otype", "all" } };
    if ( config.contains( "antennas" ) )
        antennas = config[ "antennas" ];
    const cli::AntennaPattern pattern = cli::antennaPattern( antennas, std::max<uint32_t>( 16, (uint32_t)rxTrain.size() ) );
    rxTrain = cli::selectAntennas( rxTrain, pattern );
    rxData = cli::selectAntennas( rxData, pattern );
    if ( rxTrain.empty() )
        throw std::invalid_argument( "no antenna selected" );
    const uint32_t numAntennas = (uint32_t)rxTrain.size();
    if ( user >= txTrain.size() )
        throw std::invalid_argument( "user " + std::to_string( user ) + " is not in the reference data" );
    for ( auto* v : { &txTrain, &txData } )
        for ( auto& row : *v )
            if ( row.size() < ( v == &txTrain ? rxTrain[ 0 ].size() : rxData[ 0 ].size() ) )
                throw std::invalid_argument( "reference data holds fewer symbols than the received signal" );
    // the reference symbols may be longer than the received signal: use the leading part
    for ( auto& row : txTrain )
        row.resize( rxTrain[ 0 ].size() );
    for ( auto& row : txData )
        row.resize( rxData[ 0 ].size() );

    const nlohmann::json algorithm = config[ "algorithm" ];
    if ( args.flag( "--parallel-processing" ) )
    {
        // the reference tool detects its default user count (cpp/cli/common/util.hpp:103); the
        // bundled files carry two more, all-zero user columns
        const uint32_t U = std::min<uint32_t>( cli::kDefaultUsers, (uint32_t)txTrain.size() );
        ComplexSampleMatrix est( U );
        if ( algorithm[ "otype" ] == "APSM" )
        {
            auto det = ApsmNomaDetectorFactory::build( algorithm, numAntennas, U );
            auto* apsm = dynamic_cast<ApsmDetectorBase*>( det.get() );
            apsm->train( rxTrain, ComplexSampleMatrix( txTrain.begin(), txTrain.begin() + U ) );
            apsm->detect( rxData, est );
        }
        else
        {
            for ( uint32_t u = 0; u < U; u++ )
            {
                auto det = NomaDetectorFactory::build( algorithm, numAntennas );
                det->train( rxTrain, txTrain[ u ] );
                det->detect( rxData, est[ u ] );
            }
        }
        for ( uint32_t u = 0; u < U; u++ )
            cli::printErrors( mode, est[ u ], txData[ u ], mod );
        return EXIT_SUCCESS;
    }

    const std::string historyFile = args.get( "--train-history-file" );
    NomaTrainHistory stats( mod.getBitPerSymbol(), mod.getScale() );
    NomaValSet valSet = std::nullopt;
    if ( !historyFile.empty() )
        valSet.emplace( std::make_tuple( std::ref( stats ), std::cref( rxData ), std::cref( txData[ user ] ) ) );

    ComplexSampleVector est;
    std::unique_ptr<NomaDetector> d = NomaDetectorFactory::build( algorithm, numAntennas );
    const float trainTime = d->train( rxTrain, txTrain[ user ], valSet );
    const float detectTime = d->detect( rxData, est );
    std::cout << "Train time = " << trainTime << " ms, detect time = " << detectTime << " ms" << std::endl;
    cli::printErrors( mode, est, txData[ user ], mod );

    if ( args.present( "--detected-data" ) )
        cli::writeBinary( args.get( "--detected-data" ), { { "estSigData", cli::toArray( ComplexSampleMatrix { est } ) } } );
    if ( valSet.has_value() )
        stats.toCsv( historyFile );
    return EXIT_SUCCESS;
}

} // namespace

int main( int argc, char** argv )
{
    try
    {
        return run( argc, argv );
    }
    catch ( const std::exception& e )
    {
        std::cout.clear();
        std::cerr << e.what() << std::endl;
        return EXIT_FAILURE;
    }
}
